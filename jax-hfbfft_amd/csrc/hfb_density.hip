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
// Density partials (double) go per chunk and axis into slots of
// `dout` = [DENS_G chunk][2 iso][17 slot][nvox]:
//   5 tau_x 6 j_x 7 Jy(x) 8 Jz(x)                              (x pass)
//   9 tau_y 10 j_y 11 Jx(y) 12 Jz(y)  0 rho 1 chi                (y pass)
//   13 tau_z 14 j_z 15 Jx(z) 16 Jy(z)  2 sx 3 sy 4 sz            (z pass)
// and combined by hfb_density_finish (with the linear mixing of
// static.py:1817-1823 when requested).  Per-state sums go to
// spart = [3 axis][S][ntile_a][8]: kin, Lx, Ly, Lz, sx, sy, sz, parity.
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
__device__ __forceinline__ void dens_finish_body(const double* __restrict__ d, int64_t nvox, double* rho,
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
      for (int gch = 0; gch < DENS_G; ++gch) acc += d[(((int64_t)gch * 2 + iq) * NSLOT + k) * nvox + v];
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

__global__ void dens_finish_kernel(const double* __restrict__ d, int64_t nvox, double* rho, double* chi,
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

__global__ void __launch_bounds__(256) dens_sp_finish_kernel(const double* __restrict__ d, int64_t nvox, double* rho,
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

extern "C" int hfb_density_finish(hfb_ctx* c, const double* dout, double* rho, double* chi, double* tau,
                                  double* current, double* sdens, double* sodens, int mix, double addnew,
                                  void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!dout || !rho || !chi || !tau || !current || !sdens || !sodens) return HFB_E_ARG;
  const int64_t nvox = (int64_t)c->nx * c->ny * c->nz;
  const int nb = (int)((2 * nvox + 255) / 256);
  hipLaunchKernelGGL(dens_finish_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, dout, nvox, rho, chi, tau,
                     current, sdens, sodens, mix, addnew);
  return launch_status();
}

extern "C" int hfb_density_sp_finish(hfb_ctx* c, const double* dout, double* rho, double* chi, double* tau,
                                     double* current, double* sdens, double* sodens, int mix, double addnew,
                                     const double* spart, int nst0, int nst1, double h2m0, double h2m1,
                                     double* sp_kinetic, double* sp_orbital, double* sp_spin, double* sp_parity,
                                     void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!dout || !rho || !chi || !tau || !current || !sdens || !sodens) return HFB_E_ARG;
  if (!spart || !sp_kinetic || !sp_orbital || !sp_spin || !sp_parity) return HFB_E_ARG;
  const int ns = nst0 + nst1;
  if (ns <= 0) return HFB_E_ARG;
  const int64_t nvox = (int64_t)c->nx * c->ny * c->nz;
  const int nfin = (int)((2 * nvox + 255) / 256);
  SpFin f{spart, ns, nst0, pencils_axis(0, c->nx, c->ny, c->nz) / DENS_PW,
          pencils_axis(1, c->nx, c->ny, c->nz) / DENS_PW, pencils_axis(2, c->nx, c->ny, c->nz) / DENS_PW,
          c->dx * c->dy * c->dz, h2m0, h2m1, sp_kinetic, sp_orbital, sp_spin, sp_parity};
  hipLaunchKernelGGL(dens_sp_finish_kernel, dim3(nfin + ns), dim3(256), 0, (hipStream_t)stream, dout, nvox, rho,
                     chi, tau, current, sdens, sodens, mix, addnew, nfin, f);
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
