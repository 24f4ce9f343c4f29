// Densities summed over states (densities.py:177-337) fused with the
// per-state single-particle properties (inout.py:64-149).
//
// One launch per axis a.  A workgroup owns PW pencils along a of one
// isospin and one chunk of that isospin's states; its four wavefronts take
// the chunk's states in turn (state s goes to wavefront s mod 4).  A pencil
// of both spin components sits in a group of L lanes (the LaneFFT layout of
// hfb_common.h): the spectrum gives the state's kinetic-energy contribution
// sum k^2 |psi_k|^2 / n (Nyquist kept, the Parseval form of
// inout.py:86-91,130) and the inverse transform of i k psi_k gives d_a psi
// (Nyquist zeroed, cderv*00's d1).  Density contributions accumulate in
// registers over the wavefront's states and are combined over the four
// wavefronts in LDS; the chunks are summed by hfb_density_finish, so no
// per-state field is ever written to HBM.  The pointwise fields are spread
// over the passes to even out their register load: the y pass adds rho and
// chi, the z pass the spin density, the per-state spin sums and the parity.
//
// Density partials go per chunk and axis into slots of
// `dout` = [DENS_G chunk][2 iso][17 slot][nvox]:
//   5 tau_x 6 j_x 7 Jy(x) 8 Jz(x)                              (x pass)
//   9 tau_y 10 j_y 11 Jx(y) 12 Jz(y)  0 rho 1 chi                (y pass)
//   13 tau_z 14 j_z 15 Jx(z) 16 Jy(z)  2 sx 3 sy 4 sz            (z pass)
// (doubles; the matrix-core sweep of FP32 states on 32^3 writes floats, in
// half of the scratch, and fills sz from its x pass) and combined by
// hfb_density_finish / hfb_density_sp_finish, called with the same
// precision (with the linear mixing of static.py:1817-1823 when
// requested).  Per-state sums go to spart = [3 axis][S][ntile_a][8]: kin,
// Lx, Ly, Lz, sx, sy, sz, parity.
#include <math.h>

#include "hfb_ctx.h"

using namespace hfb;

namespace {

constexpr int NSLOT = 17;

struct DensArgs {
  int nx, ny, nz;
  double dx, dy, dz;
  const void* psi;
  int nst0, nst1;  // states per isospin
  const double* wocc;
  const double* wstates;
  const double* wguv;
  const double* pairwg;
  double cm[3];
  double* dout;   // [2][17][nvox]
  double* spart;  // [S][ntile][8] for this axis
  int want_dens;
  int want_sp;
  int ntile;
};

__device__ __forceinline__ double im_ab(double ax, double ay, double bx, double by) {  // Im(a conj b)
  return ay * bx - ax * by;
}
__device__ __forceinline__ double re_ab(double ax, double ay, double bx, double by) {  // Re(a conj b)
  return ax * bx + ay * by;
}

constexpr int DENS_G = 2;  // state chunks per isospin (blocks along the states)
constexpr int DENS_L = 8;  // lanes per pencil
constexpr int DENS_PW = 64 / DENS_L;  // pencils per block

__host__ __device__ __forceinline__ int pencils_axis(int ax, int nx, int ny, int nz) {
  return ax == 0 ? ny * nz : (ax == 1 ? nx * nz : nx * ny);
}

// One workgroup of the axis pass AXIS: pencil tile `tile` of isospin iq,
// state chunk `chunk`.
template <typename T, int N, int AXIS>
__device__ __forceinline__ void dens_body(const DensArgs& a, const cx<T>* __restrict__ tw, int tile, int iq,
                                          int chunk) {
  constexpr int L = DENS_L, N1 = N / L, PW = DENS_PW;
  constexpr int NACC = AXIS == 0 ? 4 : (AXIS == 1 ? 6 : 7);
  __shared__ double comb[4][PW * N];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane % L, q = lane / L;
  const int g = tile * PW + q;  // pencil
  const int nst = iq == 0 ? a.nst0 : a.nst1;
  const int start = iq == 0 ? 0 : a.nst0;
  const int per = (nst + DENS_G - 1) / DENS_G;
  const int cs = chunk * per, ce = min(nst, cs + per);
  const int64_t nvox = (int64_t)a.nx * a.ny * a.nz;
  int64_t base, stride;
  int c0, c1;  // the two transverse coordinates of the pencil
  if constexpr (AXIS == 0) {
    base = g;
    stride = (int64_t)a.ny * a.nz;
    c0 = g / a.nz;  // y
    c1 = g % a.nz;  // z
  } else if constexpr (AXIS == 1) {
    base = (int64_t)(g / a.nz) * a.ny * a.nz + g % a.nz;
    stride = a.nz;
    c0 = g / a.nz;  // x
    c1 = g % a.nz;  // z
  } else {
    base = (int64_t)g * a.nz;
    stride = 1;
    c0 = g / a.ny;  // x
    c1 = g % a.ny;  // y
  }
  const double dax = AXIS == 0 ? a.dx : (AXIS == 1 ? a.dy : a.dz);
  const T kfac = T(2.0 * M_PI / (dax * N));
  const double kf = 2.0 * M_PI / (dax * N);
  const cx<T>* psi = (const cx<T>*)a.psi;
  LaneFFT<T, N, L> F;
  F.init(j, tw);

  double acc[N1][NACC];
#pragma unroll
  for (int i = 0; i < N1; ++i)
#pragma unroll
    for (int k = 0; k < NACC; ++k) acc[i][k] = 0.0;

#pragma unroll 1
  for (int sl = cs + wave; sl < ce; sl += 4) {
    const int s = start + sl;
    const cx<T>* p0 = psi + (int64_t)(2 * s) * nvox;
    const cx<T>* p1 = p0 + nvox;
    cx<T> u0[N1], u1[N1], g0[N1], g1[N1];
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int64_t off = base + (int64_t)(j + L * m) * stride;
      u0[m] = p0[off];
      u1[m] = p1[off];
      g0[m] = u0[m];
      g1[m] = u1[m];
    }
    F.fwd(g0, tw);
    F.fwd(g1, tw);
    double sp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int kq = F.freq(k1);
      const double mk = double(kq <= N / 2 ? kq : N - kq) * kf;
      sp[0] += mk * mk *
               (double(g0[k1].x) * double(g0[k1].x) + double(g0[k1].y) * double(g0[k1].y) +
                double(g1[k1].x) * double(g1[k1].x) + double(g1[k1].y) * double(g1[k1].y));
      const T f = d1_factor<T, N>(kq, kfac);
      g0[k1] = cmul_is(g0[k1], f);
      g1[k1] = cmul_is(g1[k1], f);
    }
    sp[0] *= 1.0 / N;
    F.inv(g0, tw);
    F.inv(g1, tw);
    const double w = a.wocc[s] * a.wstates[s];
    const double wuv = a.wguv[s] * a.pairwg[s] * a.wstates[s];
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int kk = j + L * m;
      int x, y, z;
      if constexpr (AXIS == 0) {
        x = kk; y = c0; z = c1;
      } else if constexpr (AXIS == 1) {
        x = c0; y = kk; z = c1;
      } else {
        x = c0; y = c1; z = kk;
      }
      const double xc = (x - 0.5 * (a.nx - 1)) * a.dx - a.cm[0];
      const double yc = (y - 0.5 * (a.ny - 1)) * a.dy - a.cm[1];
      const double zc = (z - 0.5 * (a.nz - 1)) * a.dz - a.cm[2];
      const double ax = u0[m].x, ay = u0[m].y, bx = u1[m].x, by = u1[m].y;
      const double cx0 = g0[m].x, cy0 = g0[m].y, cx1 = g1[m].x, cy1 = g1[m].y;
      const double tau = cx0 * cx0 + cy0 * cy0 + cx1 * cx1 + cy1 * cy1;
      const double ia = im_ab(cx0, cy0, ax, ay) + im_ab(cx1, cy1, bx, by);  // Im(d psi conj psi)
      acc[m][0] += w * tau;
      acc[m][1] += w * ia;
      const double n0 = ax * ax + ay * ay, n1 = bx * bx + by * by;
      if constexpr (AXIS == 0) {
        // densities.py:270-277
        acc[m][2] += -w * (im_ab(cx0, cy0, ax, ay) - im_ab(cx1, cy1, bx, by));   // J_y
        acc[m][3] += -w * (re_ab(ax, ay, cx1, cy1) - re_ab(bx, by, cx0, cy0));   // J_z
        sp[2] += zc * ia;
        sp[3] -= yc * ia;
      } else if constexpr (AXIS == 1) {
        // densities.py:289-296
        acc[m][2] += w * (im_ab(cx0, cy0, ax, ay) - im_ab(cx1, cy1, bx, by));    // J_x
        acc[m][3] += -w * (im_ab(cx1, cy1, ax, ay) + im_ab(cx0, cy0, bx, by));   // J_z
        acc[m][4] += w * (n0 + n1);                                                // rho
        acc[m][5] += 0.5 * wuv * (n0 + n1);                                        // chi
        sp[1] -= zc * ia;
        sp[3] += xc * ia;
      } else {
        // densities.py:308-315
        acc[m][2] += w * (re_ab(cx1, cy1, ax, ay) - re_ab(cx0, cy0, bx, by));    // J_x
        acc[m][3] += w * (im_ab(cx1, cy1, ax, ay) + im_ab(cx0, cy0, bx, by));    // J_y
        const double rsx = ax * bx + ay * by;  // Re(conj psi0 psi1)
        const double rsy = ax * by - ay * bx;  // Im(conj psi0 psi1)
        acc[m][4] += 2.0 * w * rsx;
        acc[m][5] += 2.0 * w * rsy;
        acc[m][6] += w * (n0 - n1);
        sp[1] += yc * ia;
        sp[2] -= xc * ia;
        sp[4] += 2.0 * rsx;
        sp[5] += 2.0 * rsy;
        sp[6] += n0 - n1;
        // parity: Re sum conj psi(r) psi(-r)  (inout.py:134-137)
        const int64_t mo = ((int64_t)(a.nx - 1 - x) * a.ny + (a.ny - 1 - y)) * a.nz + (a.nz - 1 - z);
        const cx<T> f0 = p0[mo], f1 = p1[mo];
        sp[7] += ax * double(f0.x) + ay * double(f0.y) + bx * double(f1.x) + by * double(f1.y);
      }
    }
    if (a.want_sp) {  // the wavefront covers this tile's pencils for state s
      // lane sums level by level over the values this pass fills (kin and
      // two L components in the x and y passes; the z pass adds the spin
      // sums and the parity, all but L_z); the others stay zero
      constexpr unsigned used = AXIS == 0 ? 0x0Du : (AXIS == 1 ? 0x0Bu : 0xF7u);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (used >> k & 1u) sp[k] += lane_xor<1>(sp[k]);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (used >> k & 1u) sp[k] += lane_xor<2>(sp[k]);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (used >> k & 1u) sp[k] += lane_xor<4>(sp[k]);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (used >> k & 1u) sp[k] += lane_xor<8>(sp[k]);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (used >> k & 1u) sp[k] += lane_xor<16>(sp[k]);
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (used >> k & 1u) sp[k] += __shfl_xor(sp[k], 32, 64);
      if (lane < 8) {
        double v = sp[0];
#pragma unroll
        for (int k = 1; k < 8; ++k)
          if (lane == k) v = sp[k];
        a.spart[((int64_t)s * a.ntile + tile) * 8 + lane] = v;
      }
    }
  }
  if (!a.want_dens) return;
  // combine the four wavefronts' accumulators in LDS, slot by slot
#pragma unroll
  for (int k = 0; k < NACC; ++k) {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < N1; ++m) comb[wave][q * N + j + L * m] = acc[m][k];
    __syncthreads();
    int slot;
    if constexpr (AXIS == 0)
      slot = 5 + k;
    else if constexpr (AXIS == 1)
      slot = k < 4 ? 9 + k : k - 4;  // rho 0, chi 1
    else
      slot = k < 4 ? 13 + k : k - 2;  // sx 2, sy 3, sz 4
    double* dst = a.dout + (((int64_t)chunk * 2 + iq) * NSLOT + slot) * nvox;
    for (int e = threadIdx.x; e < PW * N; e += 256) {
      int qq, kk;
      if constexpr (AXIS == 2) {  // contiguous along the pencil
        qq = e / N;
        kk = e % N;
      } else {  // contiguous across pencils
        qq = e % PW;
        kk = e / PW;
      }
      const int pe = qq * N + kk;
      const double v = comb[0][pe] + comb[1][pe] + comb[2][pe] + comb[3][pe];
      const int gg = tile * PW + qq;
      int64_t bb;
      if constexpr (AXIS == 0)
        bb = gg;
      else if constexpr (AXIS == 1)
        bb = (int64_t)(gg / a.nz) * a.ny * a.nz + gg % a.nz;
      else
        bb = (int64_t)gg * a.nz;
      dst[bb + (int64_t)kk * stride] = v;
    }
  }
}

template <typename T, int N, int AXIS>
__global__ void __launch_bounds__(256) dens_lane_kernel(DensArgs a, const cx<T>* __restrict__ tw) {
  dens_body<T, N, AXIS>(a, tw, blockIdx.x, blockIdx.y, blockIdx.z);
}

// All three axis passes of a cubic grid in one launch: the workgroups of
// the x, y and z passes follow each other along blockIdx.x (the passes are
// independent: each reads the states and writes its own slots).
template <typename T, int N>
__global__ void __launch_bounds__(256) dens_all_kernel(DensArgs a, const cx<T>* __restrict__ tw, int ntile,
                                                       double* sp0, double* sp1, double* sp2) {
  const int t = blockIdx.x;
  a.ntile = ntile;
  if (t < ntile) {
    a.spart = sp0;
    dens_body<T, N, 0>(a, tw, t, blockIdx.y, blockIdx.z);
  } else if (t < 2 * ntile) {
    a.spart = sp1;
    dens_body<T, N, 1>(a, tw, t - ntile, blockIdx.y, blockIdx.z);
  } else {
    a.spart = sp2;
    dens_body<T, N, 2>(a, tw, t - 2 * ntile, blockIdx.y, blockIdx.z);
  }
}

// ---------------------------------------------------------------------------
// FP32 states on 32^3 grids: the three axis passes on the matrix cores.
// Along an axis the derivative with the Nyquist mode zeroed (cderv*00's d1)
// is the dense matrix der1 of grids.py:36-50, so d1 of a tile of 32 points
// along the axis by 16 columns is one 32x32 by 32x16 product per real
// component (v_mfma_f32_16x16x4f32: 2 row blocks x 8 k-steps).  Lane
// (q, c) = (l / 16, l % 16) holds the points at positions 4 m + q, m = 0..7,
// of column c; the rows of der1 are permuted so that result register r of
// row block b is the derivative at the lane's own point m = 4 b + r, which
// feeds the density terms without any data exchange.  The kinetic energy
// keeps the Nyquist mode (inout.py:86-91): by Parseval, sum k^2 |psi_k|^2 / N
// = sum_x |d1 psi|^2 + k_N^2 |psi_N|^2 / N, psi_N the alternating sum along
// the pencil.
//
// A workgroup owns a pair of tiles of one axis, isospin and state chunk:
// wavefronts 0, 2 take the first tile, 1, 3 the second, and the wavefronts
// of a phase (0 and 1, 2 and 3) take the same states (every second state of
// the chunk).  A wavefront streams its states through two 8 KB LDS slots:
// the next state's tile goes from HBM straight into LDS (global_load_lds,
// no registers held) while the current one is on the matrix cores, so the
// load latency hides behind the work of a state.  For the x pass the pair
// is a tile and its mirror image through the origin, so the parity sums
// (psi(r) against psi(-r)) read the partner wavefront's slot; the two
// wavefronts meet at a workgroup barrier per state.  The density terms are
// those of dens_body, accumulated in FP32 per point across a wavefront's
// states and summed in FP64 across the wavefronts and chunks.  A 512-point
// tile covers two of dens_body's 256-point tiles in the per-state sums: it
// writes its sums to the first and zeros to the second.
// ---------------------------------------------------------------------------
constexpr int DM_N = 32;
constexpr int DM_PAIRS = DM_N;  // workgroups per axis: 64 tiles (32 planes x 2 column halves) in pairs
constexpr int DM_SLOT = 2 * DM_N * 16;  // complex values of one staged tile (both spins)

__device__ __forceinline__ int dm_row(int blk, int row) { return 16 * blk + 4 * (row % 4) + row / 4; }

// LDS index (complex) of the value at position p of column c of one spin in
// a staged tile: position-major for the x and y passes (columns contiguous in
// HBM), pairs of positions per column for the z pass (positions contiguous)
template <int AXIS>
__device__ __forceinline__ int dm_lidx(int spin, int p, int c) {
  if constexpr (AXIS == 2)
    return (((spin * 16 + (p >> 1)) * 16 + c) << 1) + (p & 1);
  else
    return (spin * 32 + p) * 16 + c;
}

// The tile (32 positions x 16 columns, both spins: 8 KB) of one state from
// HBM into LDS: 8 loads of 16 bytes per lane; load i fills dst[128 i, 128 i + 128).
template <int AXIS>
__device__ __forceinline__ void dm_stage(const cx<float>* p0, unsigned tbase, unsigned nvox, cx<float>* dst,
                                         int lane) {
  constexpr int N = DM_N;
  constexpr int sa = AXIS == 0 ? N * N : (AXIS == 1 ? N : 1);
  constexpr int sc = AXIS == 2 ? N : 1;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    unsigned off;
    if constexpr (AXIS == 2) {  // lane: column lane % 16, positions 2 j, 2 j + 1
      const int spin = i >> 2, j = 4 * (i & 3) + (lane >> 4);
      off = spin * nvox + tbase + (lane & 15) * sc + 2 * j;
    } else {  // lane: position row % 32, columns 2 (lane % 8), +1
      const int row = 8 * i + (lane >> 3);
      off = (row >> 5) * nvox + tbase + (row & 31) * sa + 2 * (lane & 7);
    }
    __builtin_amdgcn_global_load_lds((const void*)(p0 + off), (lds_void)(dst + 128 * i), 16, 0, 0);
  }
}

// Sum over the 64 lanes of a wavefront, returned to all lanes from lane 63
// (DPP: quad permutations, row shifts by 4 and 8, then the broadcasts of
// lane 15 and lane 31 to the following rows).
__device__ __forceinline__ float wave_total(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xF, 0xF, false));  // row_shr:4
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xF, 0xF, false));  // row_shr:8
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xA, 0xF, false));  // row_bcast:15
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x143, 0xC, 0xF, false));  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

template <int AXIS>
__device__ __forceinline__ void dm_tile(int pair, int tw, int& third, int& half) {
  if constexpr (AXIS == 0) {  // tile (y, h) and its mirror (N-1-y, 1-h), y < N/2
    const int y = pair >> 1, h = pair & 1;
    third = tw ? DM_N - 1 - y : y;
    half = tw ? 1 - h : h;
  } else {  // the two column halves of plane `pair`
    third = pair;
    half = tw;
  }
}

template <int AXIS>
__device__ __forceinline__ void densm_body(const DensArgs& a, const float* __restrict__ D1, cx<float>* buf,
                                           int pair, int iq, int chunk) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  constexpr int N = DM_N;
  // the point terms without derivatives are spread so that each pass holds
  // 5-6 accumulators: s_z and the parity sum go with the x pass
  constexpr int NACC = AXIS == 0 ? 5 : 6;
  // strides (elements) along the axis, across the columns, of the third index
  constexpr int sa = AXIS == 0 ? N * N : (AXIS == 1 ? N : 1);
  constexpr int sc = AXIS == 2 ? N : 1;
  constexpr int st = AXIS == 0 ? N : N * N;
  constexpr unsigned used = AXIS == 0 ? 0xCDu : (AXIS == 1 ? 0x0Bu : 0x37u);  // per-state sums this pass fills
  // the wavefront index as a scalar: the state index, the weights and the
  // LDS slots are then wavefront-uniform (scalar loads, no vector waits)
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tw = wave & 1, ph = wave >> 1;
  const int q = lane >> 4, c = lane & 15;
  int third, half;
  dm_tile<AXIS>(pair, tw, third, half);
  const int tile = 2 * third + half;
  const int col = half * 16 + c;
  const unsigned tbase = (unsigned)(third * st + half * 16 * sc);
  const int nst = iq == 0 ? a.nst0 : a.nst1;
  const int start = iq == 0 ? 0 : a.nst0;
  const int per = (nst + DENS_G - 1) / DENS_G;
  const int cs = chunk * per, ce = min(nst, cs + per);
  const int nit = (max(ce - cs, 0) + 1) / 2;  // the same for all four wavefronts
  const unsigned nvox = (unsigned)(N * N * N);
  const double dax = AXIS == 0 ? a.dx : (AXIS == 1 ? a.dy : a.dz);
  const double kn2 = (M_PI / dax) * (M_PI / dax) / N;  // k_N^2 / N
  const cx<float>* psi = (const cx<float>*)a.psi;
  cx<float>* const mine = buf + 2 * DM_SLOT * wave;
  const cx<float>* const partner = buf + 2 * DM_SLOT * (wave ^ 1);
  if (cs + ph < ce) dm_stage<AXIS>(psi + (size_t)(2 * (start + cs + ph)) * nvox, tbase, nvox, mine, lane);
  float Af[2][8];
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int k = 0; k < 8; ++k) Af[b][k] = D1[dm_row(b, c) * N + 4 * k + q];
  // the weights of the chunk's states, 64 at a time, one state per lane
  // (a vector load in the loop would have to wait for the staging loads
  // issued after it); a wavefront reads its state's lane with readlane
  float wl = 0.f, wuvl = 0.f;
  int wblk = 0;
  auto load_w = [&](int blk) {
    const int sx = cs + 64 * blk + lane;
    if (sx < ce) {
      const int s = start + sx;
      wl = float(a.wocc[s] * a.wstates[s]);
      wuvl = float(a.wguv[s] * a.pairwg[s] * a.wstates[s]);
    }
  };
  load_w(0);
  vm_wait<0>();  // der1, the weights and the first tile in place
  float acc[8][NACC];  // this wavefront's states; FP64 across wavefronts and chunks
#pragma unroll
  for (int m = 0; m < 8; ++m)
#pragma unroll
    for (int k = 0; k < NACC; ++k) acc[m][k] = 0.f;

#pragma unroll 1
  for (int it = 0; it < nit; ++it) {
    const int sl = cs + ph + 2 * it;
    const int slot = it & 1;
    // stage the next state into the other slot, then wait for this one:
    // loads complete in order, so at most the 8 just issued stay in flight
    if (sl + 2 < ce) {
      dm_stage<AXIS>(psi + (size_t)(2 * (start + sl + 2)) * nvox, tbase, nvox, mine + DM_SLOT * (slot ^ 1), lane);
      vm_wait<8>();
    } else {
      vm_wait<0>();
    }
    if constexpr (AXIS == 0) __syncthreads();  // the partner's tile is in LDS
    if (sl < ce) {
      const int s = start + sl;
      const cx<float>* tb = mine + DM_SLOT * slot;
      float ar[8], ai[8], br[8], bi[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const cx<float> va = tb[dm_lidx<AXIS>(0, 4 * m + q, c)], vb = tb[dm_lidx<AXIS>(1, 4 * m + q, c)];
        ar[m] = va.x;
        ai[m] = va.y;
        br[m] = vb.x;
        bi[m] = vb.y;
      }
      const int rel = sl - cs;
      if ((rel >> 6) != wblk) {  // chunks of more than 64 states
        wblk = rel >> 6;
        load_w(wblk);
      }
      const float w = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wl), rel & 63));
      const float wuv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wuvl), rel & 63));
      // the point terms in FP32 (products of FP32 values), per-state sums in
      // FP32 over the lane's 8 points; one 16-row block of d1 at a time
      float spf[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        v4f t0 = {0.f, 0.f, 0.f, 0.f}, t1 = t0, t2 = t0, t3 = t0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          t0 = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[b][k], ar[k], t0, 0, 0, 0);
          t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[b][k], ai[k], t1, 0, 0, 0);
          t2 = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[b][k], br[k], t2, 0, 0, 0);
          t3 = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[b][k], bi[k], t3, 0, 0, 0);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = 4 * b + r;
          const int pos = 4 * m + q;
          int x, y, z;
          if constexpr (AXIS == 0) {
            x = pos; y = third; z = col;
          } else if constexpr (AXIS == 1) {
            x = third; y = pos; z = col;
          } else {
            x = third; y = col; z = pos;
          }
          const float xc = float((x - 0.5 * (a.nx - 1)) * a.dx - a.cm[0]);
          const float yc = float((y - 0.5 * (a.ny - 1)) * a.dy - a.cm[1]);
          const float zc = float((z - 0.5 * (a.nz - 1)) * a.dz - a.cm[2]);
          const float ax = ar[m], ay = ai[m], bx = br[m], by = bi[m];
          const float cx0 = t0[r], cy0 = t1[r], cx1 = t2[r], cy1 = t3[r];
          const float tau = cx0 * cx0 + cy0 * cy0 + cx1 * cx1 + cy1 * cy1;
          const float i0 = cy0 * ax - cx0 * ay, i1 = cy1 * bx - cx1 * by;  // Im(d psi conj psi) per spin
          const float ia = i0 + i1;
          acc[m][0] += w * tau;
          acc[m][1] += w * ia;
          spf[0] += tau;
          const float n0 = ax * ax + ay * ay, n1 = bx * bx + by * by;
          if constexpr (AXIS == 0) {
            // densities.py:270-277
            acc[m][2] += -w * (i0 - i1);                                                  // J_y
            acc[m][3] += -w * ((ax * cx1 + ay * cy1) - (bx * cx0 + by * cy0));           // J_z
            acc[m][4] += w * (n0 - n1);                                                   // s_z
            spf[2] += zc * ia;
            spf[3] -= yc * ia;
            spf[6] += n0 - n1;
            // parity: Re sum conj psi(r) psi(-r) (inout.py:134-137); -r is
            // position N-1-pos, column 15-c of the partner tile
            const cx<float>* pb = partner + DM_SLOT * slot;
            const cx<float> f0 = pb[dm_lidx<0>(0, N - 1 - pos, 15 - c)], f1 = pb[dm_lidx<0>(1, N - 1 - pos, 15 - c)];
            spf[7] += ax * f0.x + ay * f0.y + bx * f1.x + by * f1.y;
          } else if constexpr (AXIS == 1) {
            // densities.py:289-296
            acc[m][2] += w * (i0 - i1);                                                   // J_x
            acc[m][3] += -w * ((cy1 * ax - cx1 * ay) + (cy0 * bx - cx0 * by));           // J_z
            acc[m][4] += w * (n0 + n1);                                                   // rho
            acc[m][5] += 0.5f * wuv * (n0 + n1);                                          // chi
            spf[1] -= zc * ia;
            spf[3] += xc * ia;
          } else {
            // densities.py:308-315
            acc[m][2] += w * ((cx1 * ax + cy1 * ay) - (cx0 * bx + cy0 * by));            // J_x
            acc[m][3] += w * ((cy1 * ax - cx1 * ay) + (cy0 * bx - cx0 * by));            // J_y
            const float rsx = ax * bx + ay * by;  // Re(conj psi0 psi1)
            const float rsy = ax * by - ay * bx;  // Im(conj psi0 psi1)
            acc[m][4] += 2.f * w * rsx;
            acc[m][5] += 2.f * w * rsy;
            spf[1] += yc * ia;
            spf[2] -= xc * ia;
            spf[4] += 2.f * rsx;
            spf[5] += 2.f * rsy;
          }
        }
      }
      // psi_N of the lane's column: (-1)^position = (-1)^q on this lane
      float nar = 0.f, nai = 0.f, nbr = 0.f, nbi = 0.f;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        nar += ar[m];
        nai += ai[m];
        nbr += br[m];
        nbi += bi[m];
      }
      if (a.want_sp) {
        // the Nyquist term once per column: psi_N = sum over the four q lanes
        // (lanes l ^ 16, l ^ 32) of (-1)^q times their sums
        const float sg = (q & 1) ? -1.f : 1.f;
        nar *= sg;
        nai *= sg;
        nbr *= sg;
        nbi *= sg;
        nar += __shfl_xor(nar, 16, 64);
        nai += __shfl_xor(nai, 16, 64);
        nbr += __shfl_xor(nbr, 16, 64);
        nbi += __shfl_xor(nbi, 16, 64);
        nar += __shfl_xor(nar, 32, 64);
        nai += __shfl_xor(nai, 32, 64);
        nbr += __shfl_xor(nbr, 32, 64);
        nbi += __shfl_xor(nbi, 32, 64);
        if (q == 0) spf[0] += float(kn2) * (nar * nar + nai * nai + nbr * nbr + nbi * nbi);
        // the tile's totals in FP32 (a summation tree of depth 9 over the
        // 512 points: relative error below 1e-6 of the summed magnitudes),
        // FP64 across the tiles
        float tot[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) tot[k] = (used >> k & 1u) ? wave_total(spf[k]) : 0.f;
        if (lane < 16) {
          float v = 0.f;
#pragma unroll
          for (int k = 0; k < 8; ++k)
            if (lane == k) v = tot[k];
          a.spart[((int64_t)s * a.ntile + 2 * tile) * 8 + lane] = v;  // lanes 8..15: the second, empty slot
        }
      }
    }
    if constexpr (AXIS == 0) __syncthreads();  // both tiles read before the slot is staged again
  }
  if (!a.want_dens) return;
  // all accumulators of the four wavefronts to LDS (the staging slots, idle
  // now; every load into them has completed) in one round, then the sum of
  // each tile's two wavefronts to the FP32 partial slots (the chunks are
  // summed in FP64 by the finish)
  float* comb = reinterpret_cast<float*>(buf);
  auto pidx = [&](int pos, int cc) { return AXIS == 2 ? cc * N + pos : pos * 16 + cc; };
  int thirds[2], halves[2];
  dm_tile<AXIS>(pair, 0, thirds[0], halves[0]);
  dm_tile<AXIS>(pair, 1, thirds[1], halves[1]);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NACC; ++k)
#pragma unroll
    for (int m = 0; m < 8; ++m) comb[(wave * NACC + k) * 16 * N + pidx(4 * m + q, c)] = acc[m][k];
  __syncthreads();
  float* const d0 = reinterpret_cast<float*>(a.dout) + ((int64_t)chunk * 2 + iq) * NSLOT * nvox;
  for (int e = threadIdx.x; e < NACC * 2 * 16 * N; e += 256) {
    const int k = e / (2 * 16 * N), t2 = (e / (16 * N)) & 1, el = e % (16 * N);
    int slot;
    if constexpr (AXIS == 0)
      slot = k < 4 ? 5 + k : 4;  // sz 4
    else if constexpr (AXIS == 1)
      slot = k < 4 ? 9 + k : k - 4;  // rho 0, chi 1
    else
      slot = k < 4 ? 13 + k : k - 2;  // sx 2, sy 3
    int pos, cc;
    if constexpr (AXIS == 2) {  // contiguous along the axis
      cc = el / N;
      pos = el % N;
    } else {  // contiguous across the columns
      pos = el / 16;
      cc = el % 16;
    }
    const int pe = pidx(pos, cc);
    const double v = double(comb[(t2 * NACC + k) * 16 * N + pe]) + double(comb[((t2 + 2) * NACC + k) * 16 * N + pe]);
    d0[(int64_t)slot * nvox + thirds[t2] * st + (halves[t2] * 16 + cc) * sc + pos * sa] = (float)v;
  }
}

// All three axis passes in one launch (blockIdx.x = axis * DM_PAIRS + pair).
__global__ void __launch_bounds__(256, 2) dens_mfma_kernel(DensArgs a, const float* __restrict__ D1x,
                                                           const float* __restrict__ D1y,
                                                           const float* __restrict__ D1z, double* sp0, double* sp1,
                                                           double* sp2) {
  __shared__ __attribute__((aligned(16))) cx<float> buf[4 * 2 * DM_SLOT];  // two staging slots per wavefront
  const int t = blockIdx.x;
  if (t < DM_PAIRS) {
    a.spart = sp0;
    densm_body<0>(a, D1x, buf, t, blockIdx.y, blockIdx.z);
  } else if (t < 2 * DM_PAIRS) {
    a.spart = sp1;
    densm_body<1>(a, D1y, buf, t - DM_PAIRS, blockIdx.y, blockIdx.z);
  } else {
    a.spart = sp2;
    densm_body<2>(a, D1z, buf, t - 2 * DM_PAIRS, blockIdx.y, blockIdx.z);
  }
}

// FP32 states on 32^3 grids with the FP32 derivative matrices of the context
bool dens_mfma_path(const hfb_ctx* c, int prec) {
  return prec == HFB_F32 && c->nx == DM_N && c->ny == DM_N && c->nz == DM_N && c->der1f[0] && c->der1f[1] &&
         c->der1f[2];
}

// The partial slots in dout are FP32 (half of the scratch) where the
// matrix-core sweep writes them, FP64 otherwise.
bool dens_partials_f32(const hfb_ctx* c, int prec) { return dens_mfma_path(c, prec); }

int launch_dens_mfma(const hfb_ctx* c, DensArgs a, double* sp0, double* sp1, double* sp2, hipStream_t st) {
  // the per-state sums keep dens_body's layout: 128 tile slots per axis
  a.ntile = pencils_axis(0, c->nx, c->ny, c->nz) / DENS_PW;
  static_assert(4 * DM_PAIRS == DM_N * DM_N / DENS_PW, "a tile covers two slots of the per-state sums");
  hipLaunchKernelGGL(dens_mfma_kernel, dim3(3 * DM_PAIRS, 2, DENS_G), dim3(256), 0, st, a, c->der1f[0], c->der1f[1],
                     c->der1f[2], sp0, sp1, sp2);
  return launch_status();
}

template <typename T>
int launch_dens_all(const hfb_ctx* c, DensArgs a, double* sp0, double* sp1, double* sp2, hipStream_t st) {
  const int ntile = pencils_axis(0, c->nx, c->ny, c->nz) / DENS_PW;
  HFB_DISPATCH_N(c->nx, {
    hipLaunchKernelGGL((dens_all_kernel<T, N_>), dim3(3 * ntile, 2, DENS_G), dim3(256), 0, st, a,
                       twiddles<T>(c, N_), ntile, sp0, sp1, sp2);
  })
  return launch_status();
}

template <typename T, int AXIS>
int launch_dens(const hfb_ctx* c, DensArgs a, hipStream_t st) {
  const int n = axis_len(c, AXIS);
  const int np = pencils_axis(AXIS, c->nx, c->ny, c->nz);
  if (np % DENS_PW != 0) return HFB_E_SIZE;
  a.ntile = np / DENS_PW;
  HFB_DISPATCH_N(n, {
    hipLaunchKernelGGL((dens_lane_kernel<T, N_, AXIS>), dim3(a.ntile, 2, DENS_G), dim3(256), 0, st, a,
                       twiddles<T>(c, N_));
  })
  return launch_status();
}

// Sum the axis slots, optional mixing new*addnew + old*(1-addnew) of rho,
// tau, chi (static.py:1817-1823).
// (workgroup vb of vg workgroups)
// P: the type of the partial slots (FP32 from the matrix-core sweep, else FP64)
template <typename P>
__device__ __forceinline__ void dens_finish_body(const P* __restrict__ d, int64_t nvox, double* rho,
                                                 double* chi, double* tau, double* cur, double* sdens,
                                                 double* sodens, int mix, double addnew, int vb, int vg) {
  const int64_t tot = 2 * nvox;
  for (int64_t e = vb * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)vg * blockDim.x) {
    const int iq = e / nvox;
    const int64_t v = e % nvox;
    double sl[NSLOT];
#pragma unroll
    for (int k = 0; k < NSLOT; ++k) {
      double acc = 0.0;
      for (int gch = 0; gch < DENS_G; ++gch) acc += (double)d[(((int64_t)gch * 2 + iq) * NSLOT + k) * nvox + v];
      sl[k] = acc;
    }
    const double* s = sl;
    double r = s[0], ch = s[1];
    double ta = s[5] + s[9] + s[13];
    if (mix) {
      const double ac = 1.0 - addnew;
      r = addnew * r + ac * rho[e];
      ta = addnew * ta + ac * tau[e];
      ch = addnew * ch + ac * chi[e];
    }
    rho[e] = r;
    chi[e] = ch;
    tau[e] = ta;
    const int64_t b3 = (int64_t)iq * 3 * nvox + v;
    cur[b3] = s[6];
    cur[b3 + nvox] = s[10];
    cur[b3 + 2 * nvox] = s[14];
    sdens[b3] = s[2];
    sdens[b3 + nvox] = s[3];
    sdens[b3 + 2 * nvox] = s[4];
    sodens[b3] = s[11] + s[15];
    sodens[b3 + nvox] = s[7] + s[16];
    sodens[b3 + 2 * nvox] = s[8] + s[12];
  }
}

template <typename P>
__global__ void dens_finish_kernel(const P* __restrict__ d, int64_t nvox, double* rho, double* chi,
                                   double* tau, double* cur, double* sdens, double* sodens, int mix, double addnew) {
  dens_finish_body(d, nvox, rho, chi, tau, cur, sdens, sodens, mix, addnew, blockIdx.x, gridDim.x);
}

// Per-state properties from the axis partials (inout.py:98-148).
__device__ __forceinline__ void sp_finish_body(const double* __restrict__ spart, int ns, int nst0, int nt0, int nt1,
                                               int nt2, double wxyz, double h2m0, double h2m1, double* kin, double* orb,
                                               double* spin, double* par, int s) {
  __shared__ double red[4 * 8];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int nts[3] = {nt0, nt1, nt2};
  int64_t off = 0;
  for (int ax = 0; ax < 3; ++ax) {
    for (int tt = threadIdx.x; tt < nts[ax]; tt += blockDim.x) {
      const double* p = spart + off + ((int64_t)s * nts[ax] + tt) * 8;
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] += p[j];
    }
    off += (int64_t)ns * nts[ax] * 8;
  }
  block_sum<8>(v, red);
  if (threadIdx.x == 0) {
    const double h2m = s < nst0 ? h2m0 : h2m1;
    kin[s] = wxyz * h2m * v[0];
    orb[3 * s + 0] = wxyz * v[1];
    orb[3 * s + 1] = wxyz * v[2];
    orb[3 * s + 2] = wxyz * v[3];
    spin[3 * s + 0] = wxyz * 0.5 * v[4];
    spin[3 * s + 1] = wxyz * 0.5 * v[5];
    spin[3 * s + 2] = wxyz * 0.5 * v[6];
    par[s] = wxyz * v[7];
  }
}

__global__ void sp_finish_kernel(const double* __restrict__ spart, int ns, int nst0, int nt0, int nt1, int nt2,
                                 double wxyz, double h2m0, double h2m1, double* kin, double* orb, double* spin,
                                 double* par) {
  sp_finish_body(spart, ns, nst0, nt0, nt1, nt2, wxyz, h2m0, h2m1, kin, orb, spin, par, blockIdx.x);
}

// density finish (workgroups [0, nfin)) and sp_finish (one workgroup per
// state after them) in one launch: both only read the axis passes' output
struct SpFin {
  const double* spart;
  int ns, nst0, nt0, nt1, nt2;
  double wxyz, h2m0, h2m1;
  double *kin, *orb, *spin, *par;
};

template <typename P>
__global__ void __launch_bounds__(256) dens_sp_finish_kernel(const P* __restrict__ d, int64_t nvox, double* rho,
                                                             double* chi, double* tau, double* cur, double* sdens,
                                                             double* sodens, int mix, double addnew, int nfin,
                                                             SpFin f) {
  if ((int)blockIdx.x < nfin)
    dens_finish_body(d, nvox, rho, chi, tau, cur, sdens, sodens, mix, addnew, blockIdx.x, nfin);
  else
    sp_finish_body(f.spart, f.ns, f.nst0, f.nt0, f.nt1, f.nt2, f.wxyz, f.h2m0, f.h2m1, f.kin, f.orb, f.spin, f.par,
                   blockIdx.x - nfin);
}

}  // namespace

extern "C" int hfb_density_sizes(hfb_ctx* c, int prec, long long* dout_doubles, long long* spart_doubles_per_state) {
  if (!c) return HFB_E_NOCTX;
  const long long nvox = (long long)c->nx * c->ny * c->nz;
  if (dout_doubles) *dout_doubles = (long long)DENS_G * 2LL * NSLOT * nvox;
  if (spart_doubles_per_state)
    *spart_doubles_per_state = 8LL * (pencils_axis(0, c->nx, c->ny, c->nz) + pencils_axis(1, c->nx, c->ny, c->nz) +
                                      pencils_axis(2, c->nx, c->ny, c->nz)) / DENS_PW;
  (void)prec;
  return HFB_OK;
}

// want_dens: write the density slots to dout; want_sp: per-state sums.
extern "C" int hfb_density(hfb_ctx* c, int prec, const void* psi, int nst0, int nst1, const double* wocc,
                           const double* wstates, const double* wguv, const double* pairwg, const double* cmtot,
                           int want_dens, int want_sp, double* dout, double* spart, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!psi || !wocc || !wstates || !wguv || !pairwg || nst0 < 0 || nst1 < 0) return HFB_E_ARG;
  if ((want_dens && !dout) || (want_sp && !spart)) return HFB_E_ARG;
  DensArgs a{};
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.dx = c->dx;
  a.dy = c->dy;
  a.dz = c->dz;
  a.psi = psi;
  a.nst0 = nst0;
  a.nst1 = nst1;
  a.wocc = wocc;
  a.wstates = wstates;
  a.wguv = wguv;
  a.pairwg = pairwg;
  a.cm[0] = cmtot ? cmtot[0] : 0.0;
  a.cm[1] = cmtot ? cmtot[1] : 0.0;
  a.cm[2] = cmtot ? cmtot[2] : 0.0;
  a.dout = dout;
  a.want_dens = want_dens;
  a.want_sp = want_sp;
  const int ns = nst0 + nst1;
  const int nt0 = pencils_axis(0, c->nx, c->ny, c->nz) / DENS_PW;
  const int nt1 = pencils_axis(1, c->nx, c->ny, c->nz) / DENS_PW;
  hipStream_t st = (hipStream_t)stream;
  int rc = HFB_OK;
  if (c->nx == c->ny && c->ny == c->nz && pencils_axis(0, c->nx, c->ny, c->nz) % DENS_PW == 0) {
    double* sp0 = spart;
    double* sp1 = spart ? spart + (int64_t)ns * nt0 * 8 : nullptr;
    double* sp2 = spart ? spart + (int64_t)ns * (nt0 + nt1) * 8 : nullptr;
    if (prec == HFB_F64) return launch_dens_all<double>(c, a, sp0, sp1, sp2, st);
    if (dens_mfma_path(c, prec)) return launch_dens_mfma(c, a, sp0, sp1, sp2, st);
    if (prec == HFB_F32) return launch_dens_all<float>(c, a, sp0, sp1, sp2, st);
    return HFB_E_ARG;
  }
  for (int ax = 0; ax < 3 && rc == HFB_OK; ++ax) {
    a.spart = spart ? spart + (ax == 0 ? 0 : (ax == 1 ? (int64_t)ns * nt0 * 8 : (int64_t)ns * (nt0 + nt1) * 8)) : nullptr;
    if (prec == HFB_F64) {
      if (ax == 0) rc = launch_dens<double, 0>(c, a, st);
      if (ax == 1) rc = launch_dens<double, 1>(c, a, st);
      if (ax == 2) rc = launch_dens<double, 2>(c, a, st);
    } else if (prec == HFB_F32) {
      if (ax == 0) rc = launch_dens<float, 0>(c, a, st);
      if (ax == 1) rc = launch_dens<float, 1>(c, a, st);
      if (ax == 2) rc = launch_dens<float, 2>(c, a, st);
    } else {
      return HFB_E_ARG;
    }
  }
  return rc;
}

extern "C" int hfb_density_finish(hfb_ctx* c, int prec, const double* dout, double* rho, double* chi,
                                  double* tau, double* current, double* sdens, double* sodens, int mix,
                                  double addnew, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!dout || !rho || !chi || !tau || !current || !sdens || !sodens) return HFB_E_ARG;
  if (prec != HFB_F32 && prec != HFB_F64) return HFB_E_ARG;
  const int64_t nvox = (int64_t)c->nx * c->ny * c->nz;
  const int nb = (int)((2 * nvox + 255) / 256);
  if (dens_partials_f32(c, prec))
    hipLaunchKernelGGL(dens_finish_kernel<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream,
                       (const float*)dout, nvox, rho, chi, tau, current, sdens, sodens, mix, addnew);
  else
    hipLaunchKernelGGL(dens_finish_kernel<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream, dout, nvox, rho,
                       chi, tau, current, sdens, sodens, mix, addnew);
  return launch_status();
}

extern "C" int hfb_density_sp_finish(hfb_ctx* c, int prec, const double* dout, double* rho, double* chi,
                                     double* tau, double* current, double* sdens, double* sodens, int mix,
                                     double addnew, const double* spart, int nst0, int nst1, double h2m0,
                                     double h2m1, double* sp_kinetic, double* sp_orbital, double* sp_spin,
                                     double* sp_parity, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!dout || !rho || !chi || !tau || !current || !sdens || !sodens) return HFB_E_ARG;
  if (prec != HFB_F32 && prec != HFB_F64) return HFB_E_ARG;
  if (!spart || !sp_kinetic || !sp_orbital || !sp_spin || !sp_parity) return HFB_E_ARG;
  const int ns = nst0 + nst1;
  if (ns <= 0) return HFB_E_ARG;
  const int64_t nvox = (int64_t)c->nx * c->ny * c->nz;
  const int nfin = (int)((2 * nvox + 255) / 256);
  SpFin f{spart, ns, nst0, pencils_axis(0, c->nx, c->ny, c->nz) / DENS_PW,
          pencils_axis(1, c->nx, c->ny, c->nz) / DENS_PW, pencils_axis(2, c->nx, c->ny, c->nz) / DENS_PW,
          c->dx * c->dy * c->dz, h2m0, h2m1, sp_kinetic, sp_orbital, sp_spin, sp_parity};
  if (dens_partials_f32(c, prec))
    hipLaunchKernelGGL(dens_sp_finish_kernel<float>, dim3(nfin + ns), dim3(256), 0, (hipStream_t)stream,
                       (const float*)dout, nvox, rho, chi, tau, current, sdens, sodens, mix, addnew, nfin, f);
  else
    hipLaunchKernelGGL(dens_sp_finish_kernel<double>, dim3(nfin + ns), dim3(256), 0, (hipStream_t)stream, dout,
                       nvox, rho, chi, tau, current, sdens, sodens, mix, addnew, nfin, f);
  return launch_status();
}

extern "C" int hfb_sp_finish(hfb_ctx* c, const double* spart, int nst0, int nst1, double h2m0, double h2m1,
                             double* sp_kinetic, double* sp_orbital, double* sp_spin, double* sp_parity,
                             void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!spart || !sp_kinetic || !sp_orbital || !sp_spin || !sp_parity) return HFB_E_ARG;
  const int ns = nst0 + nst1;
  if (ns <= 0) return HFB_E_ARG;
  hipLaunchKernelGGL(sp_finish_kernel, dim3(ns), dim3(256), 0, (hipStream_t)stream, spart, ns, nst0,
                     pencils_axis(0, c->nx, c->ny, c->nz) / DENS_PW, pencils_axis(1, c->nx, c->ny, c->nz) / DENS_PW,
                     pencils_axis(2, c->nx, c->ny, c->nz) / DENS_PW, c->dx * c->dy * c->dz, h2m0, h2m1,
                     sp_kinetic, sp_orbital, sp_spin, sp_parity);
  return launch_status();
}
