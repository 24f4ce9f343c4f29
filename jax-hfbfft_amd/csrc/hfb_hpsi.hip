// One-axis spectral derivatives and the fused single-particle Hamiltonian.
//
// h psi (meanfield.py:40-159) is built in three pencil passes, one per axis.
// Each pass stages a tile of pencils of both spin components in LDS, runs
// the FFT derivatives of that axis there (cderv02's d1 and d(B d1) with the
// Nyquist rule, and cderv00 of the spin-orbit/current combination), and adds
// its terms to the accumulator.  The x pass also adds the pointwise U and
// spin-current terms; the z pass finishes with the weighting
// pout = w h psi - wuv v_pair psi and the per-tile overlap partial sums.
#include <math.h>

#include "hfb_ctx.h"
#include "hfb_jobs.h"

using namespace hfb;

namespace {

template <typename T>
__device__ __forceinline__ cx<T>* other_of(cx<T>* p, cx<T>* a, cx<T>* b) {
  return p == a ? b : a;
}

// ---------------------------------------------------------------------------
// cderv{x,y,z}{00,01,02} on independent fields (levels.py:131-395)
// ---------------------------------------------------------------------------
template <typename T, int N, int AXIS, int P, int KIND>
__global__ void __launch_bounds__(128) cderv_kernel(const cx<T>* __restrict__ in, const T* __restrict__ B,
                                                    cx<T>* __restrict__ o1, cx<T>* __restrict__ o2, int nx,
                                                    int ny, int nz, T d, const cx<T>* __restrict__ twg) {
  __shared__ cx<T> tw[N];
  __shared__ cx<T> buf[4][N * P];
  __shared__ T mb[P];
  __shared__ cx<T> nyq[P];
  const int tile = blockIdx.x;
  const int64_t nvox = (int64_t)nx * ny * nz;
  const int64_t foff = (int64_t)blockIdx.y * nvox;
  const Tile t = make_tile<AXIS>(tile, nx, ny, nz, P);
  load_twiddles(tw, twg, N);
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    int k, p;
    elem_of<AXIS, N, P>(e, k, p);
    buf[0][k * P + p] = in[foff + t.base + k * t.sk + p * t.sp];
  }
  if constexpr (KIND == 2) {
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
      T s = T(0);
      for (int k = 0; k < N; ++k) s += B[t.base + k * t.sk + p * t.sp];
      mb[p] = s / T(N);
    }
  }
  __syncthreads();
  const T kfac = T(2.0 * M_PI) / (d * T(N));
  cx<T>* r = fft_lds<T, N, P, -1>(buf[0], buf[1], tw, P);
  cx<T>* o = other_of(r, buf[0], buf[1]);
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    const int k = e / P;
    const cx<T> v = r[e];
    if constexpr (KIND == 1) buf[2][e] = cscale(v, -k2_factor<T, N>(k, kfac) / T(N));
    if constexpr (KIND == 2) {
      if (k == N / 2) nyq[e % P] = v;
    }
    r[e] = cmul_is(v, d1_factor<T, N>(k, kfac));
  }
  __syncthreads();
  cx<T>* d1 = fft_lds<T, N, P, +1>(r, o, tw, P);
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    int k, p;
    elem_of<AXIS, N, P>(e, k, p);
    const cx<T> v = d1[k * P + p];
    o1[foff + t.base + k * t.sk + p * t.sp] = v;
    if constexpr (KIND == 2) buf[2][k * P + p] = cscale(v, B[t.base + k * t.sk + p * t.sp]);
  }
  __syncthreads();
  if constexpr (KIND == 1) {
    cx<T>* d2 = fft_lds<T, N, P, +1>(buf[2], buf[3], tw, P);
    for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
      int k, p;
      elem_of<AXIS, N, P>(e, k, p);
      o2[foff + t.base + k * t.sk + p * t.sp] = d2[k * P + p];
    }
  }
  if constexpr (KIND == 2) {
    cx<T>* g = fft_lds<T, N, P, -1>(buf[2], buf[3], tw, P);
    const T nq = -T(M_PI / (double)d) * T(M_PI / (double)d) / T(N);
    for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
      const int k = e / P, p = e % P;
      if (k == N / 2)
        g[e] = cscale(nyq[p], nq * mb[p]);
      else
        g[e] = cmul_is(g[e], d1_factor<T, N>(k, kfac));
    }
    __syncthreads();
    cx<T>* d2 = fft_lds<T, N, P, +1>(g, other_of(g, buf[2], buf[3]), tw, P);
    for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
      int k, p;
      elem_of<AXIS, N, P>(e, k, p);
      o2[foff + t.base + k * t.sk + p * t.sp] = d2[k * P + p];
    }
  }
}

template <typename T, int N, int AXIS, int KIND>
int launch_cderv(const hfb_ctx* c, const void* in, const void* B, void* o1, void* o2, int nf,
                 hipStream_t st) {
  constexpr int P = 8;
  const int nt = tiles_per_field<AXIS>(c->nx, c->ny, c->nz, P);
  hipLaunchKernelGGL((cderv_kernel<T, N, AXIS, P, KIND>), dim3(nt, nf), dim3(128), 0, st,
                     (const cx<T>*)in, (const T*)B, (cx<T>*)o1, (cx<T>*)o2, c->nx, c->ny, c->nz,
                     T(axis_d(c, AXIS)), twiddles<T>(c, N));
  return launch_status();
}

template <typename T>
int cderv_dispatch(const hfb_ctx* c, int axis, int kind, const void* in, const void* B, void* o1, void* o2,
                   int nf, hipStream_t st) {
  const int n = axis_len(c, axis);
#define HFB_CD(AX)                                                                        \
  if (axis == AX) {                                                                       \
    HFB_DISPATCH_N(n, {                                                                   \
      if (kind == 0) return launch_cderv<T, N_, AX, 0>(c, in, B, o1, o2, nf, st);         \
      if (kind == 1) return launch_cderv<T, N_, AX, 1>(c, in, B, o1, o2, nf, st);         \
      return launch_cderv<T, N_, AX, 2>(c, in, B, o1, o2, nf, st);                        \
    })                                                                                    \
  }
  HFB_CD(0)
  HFB_CD(1)
  HFB_CD(2)
#undef HFB_CD
  return HFB_E_ARG;
}

// ---------------------------------------------------------------------------
// Fused h psi, one axis per launch.
// ---------------------------------------------------------------------------
template <typename T>
struct MF {
  const T* U;  // (2, nvox)
  const T* B;  // (2, nvox)
  const T* A;  // (2, 3, nvox)
  const T* S;  // (2, 3, nvox)
  const T* W;  // (2, 3, nvox)
  const T* V;  // (2, nvox)
};

// The contribution of axis AXIS to h psi on one pencil (meanfield.py:61-145,
// levels.py:283-395), in the LaneFFT layout of L lanes: a, b the two spin
// components, B, A the axis' effective mass and current potential, wa, wb
// the two spin-orbit components the axis uses (W1, W2 for x; W0, W2 for y;
// W0, W1 for z), d the grid spacing.  o0, o1 receive
//   d(f) - d(B d1) + (the axis' spin-orbit and current terms on d1),
// where d1 = IFFT(i k FFT psi) and the Nyquist entry of d(B d1) is
// -(pi/d)^2 mean(B) psi_N / N.
template <typename T, int N, int L, int AXIS>
__device__ __forceinline__ void axis_terms(const LaneFFT<T, N, L>& F, const cx<T>* __restrict__ tw,
                                           const cx<T> (&a)[N / L], const cx<T> (&b)[N / L], const T (&Bv)[N / L],
                                           const T (&Av)[N / L], const T (&wa)[N / L], const T (&wb)[N / L], T d,
                                           cx<T> (&o0)[N / L], cx<T> (&o1)[N / L]) {
  constexpr int N1 = N / L;
  cx<T> fa[N1], fb[N1], ga[N1], gb[N1];
  T bsum = T(0);
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    bsum += Bv[m];
    fa[m] = a[m];
    fb[m] = b[m];
  }
  const T mb = group_sum<L>(bsum) / T(N);
  const T kfac = T(2.0 * M_PI) / (d * T(N));
  // d1 = d/da psi
  F.fwd(fa, tw);
  F.fwd(fb, tw);
  const cx<T> nya = fa[0], nyb = fb[0];  // psi_N on the lane that holds k = N/2 at k1 = 0
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) {
    const T f = d1_factor<T, N>(F.freq(k1), kfac);
    fa[k1] = cmul_is(fa[k1], f);
    fb[k1] = cmul_is(fb[k1], f);
  }
  F.inv(fa, tw);
  F.inv(fb, tw);
  // f - B d1 with the axis' combination f (meanfield.py:75-83, 104-112, 133-141)
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const cx<T> av = a[m], bv = b[m];
    const T A = Av[m];
    cx<T> q0, q1;
    if constexpr (AXIS == 0) {  // wa = W1, wb = W2
      q0 = csub(cmul_is(av, T(-0.5) * (A - wa[m])), cscale(bv, T(0.5) * wb[m]));
      q1 = cadd(cmul_is(bv, T(-0.5) * (A + wa[m])), cscale(av, T(0.5) * wb[m]));
    } else if constexpr (AXIS == 1) {  // wa = W0, wb = W2
      q0 = cadd(cmul_is(av, T(-0.5) * (A + wa[m])), cmul_is(bv, T(0.5) * wb[m]));
      q1 = cadd(cmul_is(bv, T(-0.5) * (A - wa[m])), cmul_is(av, T(0.5) * wb[m]));
    } else {  // wa = W0, wb = W1
      q0 = cadd(cmul_is(av, T(-0.5) * A), cadd(cscale(bv, T(0.5) * wa[m]), cmul_is(bv, T(-0.5) * wb[m])));
      q1 = cadd(cmul_is(bv, T(-0.5) * A), cadd(cscale(av, T(-0.5) * wa[m]), cmul_is(av, T(-0.5) * wb[m])));
    }
    ga[m] = csub(q0, cscale(fa[m], Bv[m]));
    gb[m] = csub(q1, cscale(fb[m], Bv[m]));
  }
  F.fwd(ga, tw);
  F.fwd(gb, tw);
  const T nq = -T(M_PI / (double)d) * T(M_PI / (double)d) / T(N);
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) {
    const int kf = F.freq(k1);
    if (kf == N / 2) {  // only d(B d1) has a Nyquist entry (levels.py:309-315)
      ga[k1] = cscale(nya, -nq * mb);
      gb[k1] = cscale(nyb, -nq * mb);
    } else {
      const T f = d1_factor<T, N>(kf, kfac);
      ga[k1] = cmul_is(ga[k1], f);
      gb[k1] = cmul_is(gb[k1], f);
    }
  }
  F.inv(ga, tw);
  F.inv(gb, tw);
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const cx<T> da = fa[m], db = fb[m];
    const T A = Av[m];
    cx<T> r0 = ga[m], r1 = gb[m];
    if constexpr (AXIS == 0) {  // meanfield.py:63-73
      r0 = cadd(r0, csub(cmul_is(da, -(T(0.5) * A - T(0.5) * wa[m])), cscale(db, T(0.5) * wb[m])));
      r1 = cadd(r1, cadd(cmul_is(db, -(T(0.5) * A + T(0.5) * wa[m])), cscale(da, T(0.5) * wb[m])));
    } else if constexpr (AXIS == 1) {  // meanfield.py:92-102
      r0 = cadd(r0, cadd(cmul_is(da, -(T(0.5) * A + T(0.5) * wa[m])), cmul_is(db, T(0.5) * wb[m])));
      r1 = cadd(r1, cadd(cmul_is(db, -(T(0.5) * A - T(0.5) * wa[m])), cmul_is(da, T(0.5) * wb[m])));
    } else {  // meanfield.py:121-131
      r0 = cadd(r0, cadd(cmul_is(da, -T(0.5) * A), cadd(cscale(db, T(0.5) * wa[m]), cmul_is(db, T(-0.5) * wb[m]))));
      r1 = cadd(r1, cadd(cmul_is(db, -T(0.5) * A), cadd(cscale(da, T(-0.5) * wa[m]), cmul_is(da, T(-0.5) * wb[m]))));
    }
    o0[m] = r0;
    o1[m] = r1;
  }
}

// One axis of h psi for the pencils of one state, in registers: a group of L
// lanes holds one pencil of both spin components (LaneFFT layout), the
// block's four wavefronts hold 4 * 64 / L pencils.  Per pencil:
//   d1 = IFFT(i k FFT psi)                       (cderv00 / first half of cderv02)
//   dd = IFFT(i k FFT(f - B d1)) with the Nyquist entry -(pi/d)^2 mean(B) psi_N / N
//      = d(f) - d(B d1)                          (meanfield.py:61-145, levels.py:283-395)
// where f is the axis' spin-orbit / current combination, then the pointwise
// terms.  The x pass adds U psi and the spin-current coupling, the z pass
// the weighting w h psi - wuv v_pair psi and the overlap partial sums.
template <typename T, int N, int AXIS, int L>
__global__ void __launch_bounds__(256) hpsi_lane_kernel(const cx<T>* __restrict__ psi, cx<T>* __restrict__ pmf,
                                                        cx<T>* __restrict__ pout, cx<T>* __restrict__ pdel,
                                                        MF<T> mf, const double* __restrict__ w,
                                                        const double* __restrict__ wuv, int npsi0, int nx, int ny,
                                                        int nz, T d, const cx<T>* __restrict__ tw,
                                                        double* __restrict__ partial, WJob job) {
  // the first row (when there is a job): the step's weights, dispatched
  // first so that the energy tail before them runs beside the pass
  const int row0 = job_rows(job);
  if (row0 && blockIdx.y == 0) {
    if (blockIdx.x == 0) job_body<4>(job);
    return;
  }
  constexpr int N1 = N / L;
  constexpr int PW = 64 / L;  // pencils per wavefront
  __shared__ double red[4 * 2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane % L, q = lane / L;
  const int g = (blockIdx.x * 4 + wave) * PW + q;  // pencil
  const int s = blockIdx.y - row0;
  const int iq = s >= npsi0 ? 1 : 0;
  const int64_t nvox = (int64_t)nx * ny * nz;
  int64_t base, stride;
  if constexpr (AXIS == 0) {
    base = g;
    stride = (int64_t)ny * nz;
  } else if constexpr (AXIS == 1) {
    base = (int64_t)(g / nz) * ny * nz + g % nz;
    stride = nz;
  } else {
    base = (int64_t)g * nz;
    stride = 1;
  }
  LaneFFT<T, N, L> F;
  F.init(j, tw);
  const int64_t f0 = (int64_t)(2 * s) * nvox;
  const cx<T>* ps0 = psi + f0;
  const cx<T>* ps1 = ps0 + nvox;
  const T* Bq = mf.B + iq * nvox;
  const T* Aa = mf.A + (iq * 3 + AXIS) * nvox;
  // the two spin-orbit components this axis uses (meanfield.py:61-145)
  constexpr int WA = AXIS == 0 ? 1 : 0;
  constexpr int WB = AXIS == 2 ? 1 : 2;
  const T* Wa = mf.W + (iq * 3 + WA) * nvox;
  const T* Wb = mf.W + (iq * 3 + WB) * nvox;

  cx<T> a[N1], b[N1], o0v[N1], o1v[N1];
  T Bv[N1], Av[N1], wa[N1], wb[N1];
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const int64_t off = base + (int64_t)(j + L * m) * stride;
    a[m] = ps0[off];
    b[m] = ps1[off];
    Bv[m] = Bq[off];
    Av[m] = Aa[off];
    wa[m] = Wa[off];
    wb[m] = Wb[off];
  }
  axis_terms<T, N, L, AXIS>(F, tw, a, b, Bv, Av, wa, wb, d, o0v, o1v);

  double acc[2] = {0.0, 0.0};
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const int64_t off = base + (int64_t)(j + L * m) * stride;
    const cx<T> av = a[m], bv = b[m];
    cx<T> o0 = o0v[m], o1 = o1v[m];
    if constexpr (AXIS == 0) {  // meanfield.py:47-58: U psi and the spin-current coupling
      const T u = mf.U[iq * nvox + off];
      const T s0 = mf.S[(iq * 3 + 0) * nvox + off];
      const T s1 = mf.S[(iq * 3 + 1) * nvox + off];
      const T s2 = mf.S[(iq * 3 + 2) * nvox + off];
      o0 = cadd(o0, cadd(cscale(av, u + s2), cadd(cscale(bv, s0), cmul_is(bv, -s1))));
      o1 = cadd(o1, cadd(cscale(bv, u - s2), cadd(cscale(av, s0), cmul_is(av, s1))));
    }
    if constexpr (AXIS != 0) {
      o0 = cadd(o0, pmf[f0 + off]);
      o1 = cadd(o1, pmf[f0 + nvox + off]);
    }
    pmf[f0 + off] = o0;
    pmf[f0 + nvox + off] = o1;
    if constexpr (AXIS == 2) {  // meanfield.py:150-157
      const T v = mf.V[iq * nvox + off];
      const cx<T> e0 = cscale(av, v), e1 = cscale(bv, v);
      if (pdel) {
        pdel[f0 + off] = e0;
        pdel[f0 + nvox + off] = e1;
      }
      const T ws = T(w[s]), wu = T(wuv[s]);
      const cx<T> r0 = csub(cscale(o0, ws), cscale(e0, wu));
      const cx<T> r1 = csub(cscale(o1, ws), cscale(e1, wu));
      if (pout) {
        pout[f0 + off] = r0;
        pout[f0 + nvox + off] = r1;
      }
      acc[0] += double(av.x) * double(o0.x) + double(av.y) * double(o0.y) + double(bv.x) * double(o1.x) +
                double(bv.y) * double(o1.y);
      acc[1] += double(av.x) * double(r0.x) + double(av.y) * double(r0.y) + double(bv.x) * double(r1.x) +
                double(bv.y) * double(r1.y);
    }
  }
  if constexpr (AXIS == 2) {
    if (partial) {
      block_sum<2>(acc, red);
      if (threadIdx.x == 0) {
        partial[((int64_t)s * gridDim.x + blockIdx.x) * 2 + 0] = acc[0];
        partial[((int64_t)s * gridDim.x + blockIdx.x) * 2 + 1] = acc[1];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// FP32 x pass on the matrix cores (32-point x axis).
//
// Along an axis the spectral derivative with the Nyquist mode zeroed is the
// dense real matrix der1 (grids.py:36-50 sder; cderv00's operator,
// levels.py:131-189), so both derivatives of the x terms of hpsi01 are
// 32 x 32 matrix products over the x index, with the columns (y, z) of the
// spinor block as the other dimension:
//   d1 = D psi,  dd = D (f - B d1) + (-1)^x (pi/dx)^2 mean_x(B) psi_N / N,
// the second term being cderv02's Nyquist entry (levels.py:309-315) as a
// rank-one correction, psi_N = sum_x (-1)^x psi.  A wavefront owns 32 x
// values by 16 z values at one y of one state and multiplies the four real
// components (both spins, real and imaginary parts) with
// v_mfma_f32_16x16x4f32: the B operand of k-step s in lane l is the
// component at x = 4 s + l/16, z = l%16.  The rows of D are permuted so that
// result register r of block b in lane l holds x = 16 b + 4 r + l/16, the
// same x as operand register 4 b + r: products feed the pointwise terms
// and the second product without any data exchange.
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));

// Element at a 32-bit unsigned byte offset from a wavefront-uniform base:
// the address stays base (scalar) + offset (one vector register).
template <typename T>
__device__ __forceinline__ T ldb(const T* base, unsigned byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}
template <typename T>
__device__ __forceinline__ void stb(T* base, unsigned byte_off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// result register r of lane group q (= l/16) of a 16x16x4 product is row
// 4 q + r (hfb_selftest_mfma_f32 checks this); the position a result row holds:
__device__ __forceinline__ int mfx_row_x(int blk, int row) { return 16 * blk + 4 * (row % 4) + row / 4; }

// The axis' spin-orbit / current combination f of a two-spinor (a, b)
// (meanfield.py:75-83 x, 104-112 y, 133-141 z); the same linear form gives
// the axis terms on d1 (meanfield.py:63-73, 92-102, 121-131).  A is the
// axis' component of aq, wa/wb the two wlspot components the axis uses
// (x: W1, W2; y: W0, W2; z: W0, W1).
template <int AXIS>
__device__ __forceinline__ void fcomb(float ar, float ai, float br, float bi, float A, float wa, float wb,
                                      float& o0r, float& o0i, float& o1r, float& o1i) {
  if constexpr (AXIS == 0) {  // i p a - hw b,  i pm b + hw a
    const float p = -0.5f * (A - wa), pm = -0.5f * (A + wa), hw = 0.5f * wb;
    o0r = -ai * p - hw * br;
    o0i = ar * p - hw * bi;
    o1r = -bi * pm + hw * ar;
    o1i = br * pm + hw * ai;
  } else if constexpr (AXIS == 1) {  // i p a + i hw b,  i pm b + i hw a
    const float p = -0.5f * (A + wa), pm = -0.5f * (A - wa), hw = 0.5f * wb;
    o0r = -ai * p - bi * hw;
    o0i = ar * p + br * hw;
    o1r = -bi * pm - ai * hw;
    o1i = br * pm + ar * hw;
  } else {  // i p a + hw b + i hv b,  i p b - hw a + i hv a
    const float p = -0.5f * A, hw = 0.5f * wa, hv = -0.5f * wb;
    o0r = -ai * p + hw * br - hv * bi;
    o0i = ar * p + hw * bi + hv * br;
    o1r = -bi * p - hw * ar - hv * ai;
    o1i = br * p - hw * ai + hv * ar;
  }
}

// One wavefront's share of one axis of h psi on the matrix cores: 32
// positions along the axis by 16 columns.  Operand register m of lane l
// holds position 4 m + l/16 of column l % 16: the spinor components
// ar..bi, the effective mass Bv, the axis' aq component Av and wlspot pair
// wa, wb.  Af: the D operands (rows in result order).  extra(m, e[4]) adds
// further pointwise terms at operand position m.  Returns, in result
// register (b, r) = position 16 b + 4 r + l/16 (the operand position of
// m = 4 b + r), the axis' contribution to h psi:
//   D (f - B d1) + f(d1) + (-1)^i (pi/d)^2 mean(B) psi_N / N + extra,
// with d1 = D psi and psi_N = sum_i (-1)^i psi (levels.py:283-395).
template <int AXIS, typename Extra>
__device__ __forceinline__ void mfma_axis(const float (&Af)[2][8], const float (&ar)[8], const float (&ai)[8],
                                          const float (&br)[8], const float (&bi)[8], const float (&Bv)[8],
                                          const float (&Av)[8], const float (&wa)[8], const float (&wb)[8], float d,
                                          int q, Extra extra, v4f (&h)[4][2]) {
  constexpr int N = 32;
  v4f d1[4][2];
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
    d1[0][b] = t0;
    d1[1][b] = t1;
    d1[2][b] = t2;
    d1[3][b] = t3;
  }
  // mean(B) and psi_N over the column (its four lane groups)
  float bs = 0.f, nar = 0.f, nai = 0.f, nbr = 0.f, nbi = 0.f;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    bs += Bv[m];
    nar += ar[m];
    nai += ai[m];
    nbr += br[m];
    nbi += bi[m];
  }
  const float sg = (q & 1) ? -1.f : 1.f;  // (-1)^i, i = 4 m + q
  nar *= sg;
  nai *= sg;
  nbr *= sg;
  nbi *= sg;
#define HFB_XSUM(v) v += __shfl_xor(v, 16, 64); v += __shfl_xor(v, 32, 64);
  HFB_XSUM(bs)
  HFB_XSUM(nar)
  HFB_XSUM(nai)
  HFB_XSUM(nbr)
  HFB_XSUM(nbi)
#undef HFB_XSUM
  // cderv02's Nyquist entry (levels.py:309-315), times (-1)^i
  const float nyf = sg * float(M_PI / (double)d) * float(M_PI / (double)d) / float(N) * (bs / float(N));
  float gar[8], gai[8], gbr[8], gbi[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int b = m / 4, r = m % 4;
    const float da_r = d1[0][b][r], da_i = d1[1][b][r], db_r = d1[2][b][r], db_i = d1[3][b][r];
    float f0r, f0i, f1r, f1i;
    fcomb<AXIS>(ar[m], ai[m], br[m], bi[m], Av[m], wa[m], wb[m], f0r, f0i, f1r, f1i);
    gar[m] = f0r - Bv[m] * da_r;
    gai[m] = f0i - Bv[m] * da_i;
    gbr[m] = f1r - Bv[m] * db_r;
    gbi[m] = f1i - Bv[m] * db_i;
    float e[4];
    fcomb<AXIS>(da_r, da_i, db_r, db_i, Av[m], wa[m], wb[m], e[0], e[1], e[2], e[3]);
    e[0] += nyf * nar;
    e[1] += nyf * nai;
    e[2] += nyf * nbr;
    e[3] += nyf * nbi;
    extra(m, e);
    h[0][b][r] = e[0];
    h[1][b][r] = e[1];
    h[2][b][r] = e[2];
    h[3][b][r] = e[3];
  }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      h[0][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[b][k], gar[k], h[0][b], 0, 0, 0);
      h[1][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[b][k], gai[k], h[1][b], 0, 0, 0);
      h[2][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[b][k], gbr[k], h[2][b], 0, 0, 0);
      h[3][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[b][k], gbi[k], h[3][b], 0, 0, 0);
    }
  }
}

// The D operands of lane (q, c): D rows in result order, k = 4 kstep + q.
__device__ __forceinline__ void load_dfrag(const float* __restrict__ D1, int q, int c, float (&Af)[2][8]) {
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int k = 0; k < 8; ++k) Af[b][k] = D1[mfx_row_x(b, c) * 32 + 4 * k + q];
}

// x pass: a wavefront owns 32 x by the 16 z of one slice at one y; it adds
// U psi and the spin-current coupling (meanfield.py:47-58) and writes the
// partial h psi.
__global__ void __launch_bounds__(256, 3) hpsi_mfma_x_kernel(const cx<float>* __restrict__ psi,
                                                             cx<float>* __restrict__ pmf, MF<float> mf, int npsi0,
                                                             int ny, int nz, float d, const float* __restrict__ D1,
                                                             WJob job) {
  // the first row (when there is a job): the step's weights, dispatched
  // first so that the energy tail before them runs beside the pass
  const int row0 = job_rows(job);
  if (row0 && blockIdx.y == 0) {
    if (blockIdx.x == 0) job_body<4>(job);
    return;
  }
  constexpr int N = 32;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, c = lane & 15;
  const int nzh = nz >> 4;
  const int tile = blockIdx.x * 4 + wave;  // (y, 16-column z slice)
  const int y = tile / nzh, z = (tile % nzh) * 16 + c;
  const int s = blockIdx.y - row0;
  const int iq = s >= npsi0 ? 1 : 0;
  const int sx = ny * nz;  // offsets below: unsigned 32-bit from a uniform field base
  const int64_t nvox = (int64_t)N * sx;
  const cx<float>* pa = psi + (int64_t)(2 * s) * nvox;
  const cx<float>* pb = pa + nvox;
  const int col = y * nz + z;
  float Af[2][8];
  load_dfrag(D1, q, c, Af);
  float ar[8], ai[8], br[8], bi[8], Bv[8], Av[8], w1[8], w2[8];
  const float* Bq = mf.B + iq * nvox;
  const float* Ax = mf.A + (iq * 3 + 0) * nvox;
  const float* W1 = mf.W + (iq * 3 + 1) * nvox;
  const float* W2 = mf.W + (iq * 3 + 2) * nvox;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const unsigned off = (unsigned)((4 * m + q) * sx + col);
    const cx<float> va = ldb(pa, 8u * off), vb = ldb(pb, 8u * off);
    ar[m] = va.x;
    ai[m] = va.y;
    br[m] = vb.x;
    bi[m] = vb.y;
    Bv[m] = ldb(Bq, 4u * off);
    Av[m] = ldb(Ax, 4u * off);
    w1[m] = ldb(W1, 4u * off);
    w2[m] = ldb(W2, 4u * off);
  }
  const float* U = mf.U + iq * nvox;
  const float* S0 = mf.S + (iq * 3 + 0) * nvox;
  const float* S1 = mf.S + (iq * 3 + 1) * nvox;
  const float* S2 = mf.S + (iq * 3 + 2) * nvox;
  v4f h[4][2];
  mfma_axis<0>(Af, ar, ai, br, bi, Bv, Av, w1, w2, d, q, [&](int m, float (&e)[4]) {
    const unsigned off = (unsigned)((4 * m + q) * sx + col);
    const float u = ldb(U, 4u * off), s0 = ldb(S0, 4u * off), s1 = ldb(S1, 4u * off), s2 = ldb(S2, 4u * off);
    e[0] += ar[m] * (u + s2) + br[m] * s0 + bi[m] * s1;
    e[1] += ai[m] * (u + s2) + bi[m] * s0 - br[m] * s1;
    e[2] += br[m] * (u - s2) + ar[m] * s0 - ai[m] * s1;
    e[3] += bi[m] * (u - s2) + ai[m] * s0 + ar[m] * s1;
  }, h);
  cx<float>* oa = pmf + (int64_t)(2 * s) * nvox;
  cx<float>* ob = oa + nvox;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int b = m / 4, r = m % 4;
    const unsigned off = (unsigned)((4 * m + q) * sx + col);
    stb(oa, 8u * off, cx<float>{h[0][b][r], h[1][b][r]});
    stb(ob, 8u * off, cx<float>{h[2][b][r], h[3][b][r]});
  }
}

// The y and z axes of h psi for one y-z plane (x fixed) of one state, when
// ny == nz: the plane of both spin components is read once along z pencils
// and kept in LDS; the y pencils take it from there, compute the y terms
// and leave them in LDS; the z pencils add the z terms, the y terms and the
// x pass's accumulator, then write h psi, the weighted result and the
// overlap partial sums (one pair per plane).  N pencils of HP_L lanes.
constexpr int HP_L = 8;

template <typename T, int N>
__global__ void __launch_bounds__(N * HP_L) hpsi_plane_kernel(const cx<T>* __restrict__ psi, cx<T>* __restrict__ pmf,
                                                              cx<T>* __restrict__ pout, cx<T>* __restrict__ pdel,
                                                              MF<T> mf, const double* __restrict__ w,
                                                              const double* __restrict__ wuv, int npsi0, int nx, T d,
                                                              const cx<T>* __restrict__ tw,
                                                              double* __restrict__ partial, WJob job) {
  constexpr int L = HP_L, N1 = N / L, P = N + 1;  // P: LDS row pitch
  const int row0 = job_rows(job);
  if (row0 && blockIdx.y == 0) {  // the first row: the job (the energy tail)
    if (blockIdx.x == 0) job_body<(N * L + 63) / 64>(job);
    return;
  }
  extern __shared__ unsigned char hp_lds[];
  cx<T>* s0 = (cx<T>*)hp_lds;  // [y][z], spin up
  cx<T>* s1 = s0 + N * P;      // spin down
  __shared__ double red[(N * L / 64) * 2];
  const int j = threadIdx.x % L, q = threadIdx.x / L;
  const int x = blockIdx.x, s = blockIdx.y - row0;
  const int iq = s >= npsi0 ? 1 : 0;
  const int64_t nvox = (int64_t)nx * N * N;
  const int64_t f0 = (int64_t)(2 * s) * nvox;
  const int64_t po = (int64_t)x * N * N;
  const cx<T>* ps0 = psi + f0;
  const cx<T>* ps1 = ps0 + nvox;
  const T* Bq = mf.B + iq * nvox;
  const T* W0 = mf.W + (iq * 3 + 0) * nvox;
  const T* W1 = mf.W + (iq * 3 + 1) * nvox;
  const T* W2 = mf.W + (iq * 3 + 2) * nvox;
  LaneFFT<T, N, L> F;
  F.init(j, tw);
  // z pencils (pencil q = y): the spinor plane
  cx<T> a[N1], b[N1];
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const int z = j + L * m;
    const int64_t off = po + (int64_t)q * N + z;
    a[m] = ps0[off];
    b[m] = ps1[off];
    s0[q * P + z] = a[m];
    s1[q * P + z] = b[m];
  }
  __syncthreads();
  {  // y pencils (pencil q = z)
    cx<T> ya[N1], yb[N1], o0[N1], o1[N1];
    T Bv[N1], Av[N1], wa[N1], wb[N1];
    const T* A1 = mf.A + (iq * 3 + 1) * nvox;
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int y = j + L * m;
      const int64_t off = po + (int64_t)y * N + q;
      ya[m] = s0[y * P + q];
      yb[m] = s1[y * P + q];
      Bv[m] = Bq[off];
      Av[m] = A1[off];
      wa[m] = W0[off];
      wb[m] = W2[off];
    }
    axis_terms<T, N, L, 1>(F, tw, ya, yb, Bv, Av, wa, wb, d, o0, o1);
    __syncthreads();  // every y pencil has read the plane
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int y = j + L * m;
      s0[y * P + q] = o0[m];
      s1[y * P + q] = o1[m];
    }
  }
  __syncthreads();
  double acc[2] = {0.0, 0.0};
  {  // z pencils (pencil q = y)
    cx<T> o0[N1], o1[N1];
    T Bv[N1], Av[N1], wa[N1], wb[N1];
    const T* A2 = mf.A + (iq * 3 + 2) * nvox;
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int64_t off = po + (int64_t)q * N + j + L * m;
      Bv[m] = Bq[off];
      Av[m] = A2[off];
      wa[m] = W0[off];
      wb[m] = W1[off];
    }
    axis_terms<T, N, L, 2>(F, tw, a, b, Bv, Av, wa, wb, d, o0, o1);
    const T ws = T(w[s]), wu = T(wuv[s]);
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int z = j + L * m;
      const int64_t off = po + (int64_t)q * N + z;
      const cx<T> av = a[m], bv = b[m];
      const cx<T> h0 = cadd(cadd(o0[m], s0[q * P + z]), pmf[f0 + off]);
      const cx<T> h1 = cadd(cadd(o1[m], s1[q * P + z]), pmf[f0 + nvox + off]);
      pmf[f0 + off] = h0;
      pmf[f0 + nvox + off] = h1;
      // meanfield.py:150-157
      const T v = mf.V[iq * nvox + off];
      const cx<T> e0 = cscale(av, v), e1 = cscale(bv, v);
      if (pdel) {
        pdel[f0 + off] = e0;
        pdel[f0 + nvox + off] = e1;
      }
      const cx<T> r0 = csub(cscale(h0, ws), cscale(e0, wu));
      const cx<T> r1 = csub(cscale(h1, ws), cscale(e1, wu));
      if (pout) {
        pout[f0 + off] = r0;
        pout[f0 + nvox + off] = r1;
      }
      acc[0] += double(av.x) * double(h0.x) + double(av.y) * double(h0.y) + double(bv.x) * double(h1.x) +
                double(bv.y) * double(h1.y);
      acc[1] += double(av.x) * double(r0.x) + double(av.y) * double(r0.y) + double(bv.x) * double(r1.x) +
                double(bv.y) * double(r1.y);
    }
  }
  if (partial) {
    block_sum<2>(acc, red);
    if (threadIdx.x == 0) {
      partial[((int64_t)s * gridDim.x + x) * 2 + 0] = acc[0];
      partial[((int64_t)s * gridDim.x + x) * 2 + 1] = acc[1];
    }
  }
}

// Lanes per pencil: FP32 pencils split over 4 lanes, FP64 over 8, which
// keeps the six per-lane register arrays of the kernel at N/L elements.
template <typename T>
constexpr int lanes_per_pencil() { return 8; }

template <typename T>
int hpsi_pencils_per_block() { return 4 * (64 / lanes_per_pencil<T>()); }

inline int pencils_of_axis(const hfb_ctx* c, int axis) {
  return axis == 0 ? c->ny * c->nz : (axis == 1 ? c->nx * c->nz : c->nx * c->ny);
}

template <typename T, int AXIS>
int launch_hpsi_axis(const hfb_ctx* c, const void* psi, int ns, int npsi0, MF<T> mf, const double* w,
                     const double* wuv, void* pout, void* pmf, void* pdel, double* partial, hipStream_t st,
                     const WJob& job = WJob{}) {
  constexpr int L = lanes_per_pencil<T>();
  const int ppb = hpsi_pencils_per_block<T>();
  const int np = pencils_of_axis(c, AXIS);
  if (np % ppb != 0) return HFB_E_SIZE;
  const int n = axis_len(c, AXIS);
  const int rows = ns + job_rows(job);
  HFB_DISPATCH_N(n, {
    hipLaunchKernelGGL((hpsi_lane_kernel<T, N_, AXIS, L>), dim3(np / ppb, rows), dim3(256), 0, st,
                       (const cx<T>*)psi, (cx<T>*)pmf, (cx<T>*)pout, (cx<T>*)pdel, mf, w, wuv, npsi0, c->nx, c->ny,
                       c->nz, T(axis_d(c, AXIS)), twiddles<T>(c, N_), partial, job);
  })
  return launch_status();
}

// the plane kernel for square y-z planes; FP64 planes above 32 points per
// side would spill registers and take the pencil passes instead
inline bool hpsi_plane_path(const hfb_ctx* c, int prec) {
  return c->ny == c->nz && (prec == HFB_F32 || c->ny <= 32);
}

template <typename T>
int launch_hpsi_plane(const hfb_ctx* c, const void* psi, int ns, int npsi0, MF<T> mf, const double* w,
                      const double* wuv, void* pout, void* pmf, void* pdel, double* partial, hipStream_t st,
                      const WJob& job) {
  HFB_DISPATCH_N(c->ny, {
    constexpr int NP_ = (sizeof(T) == 8 && N_ > 32) ? 32 : N_;  // never taken for FP64 above 32
    if (NP_ != N_) return HFB_E_SIZE;
    const size_t lds = 2 * sizeof(cx<T>) * (size_t)NP_ * (NP_ + 1);
    auto kern = hpsi_plane_kernel<T, NP_>;
    if (lds > 32 * 1024) allow_full_lds((const void*)kern);
    hipLaunchKernelGGL(kern, dim3(c->nx, ns + job_rows(job)), dim3(NP_ * HP_L), lds, st, (const cx<T>*)psi,
                       (cx<T>*)pmf, (cx<T>*)pout, (cx<T>*)pdel, mf, w, wuv, npsi0, c->nx, T(axis_d(c, 1)),
                       twiddles<T>(c, NP_), partial, job);
  })
  return launch_status();
}

// FP32 spinors on a 32^3 grid: the x pass on the matrix cores, then the
// y-z plane pass (LaneFFT)
inline bool hpsi_mfma_path(const hfb_ctx* c, int prec) {
  return prec == HFB_F32 && c->nx == 32 && c->ny == 32 && c->nz == 32;
}

template <typename T>
int launch_hpsi_plane(const hfb_ctx* c, const void* psi, int ns, int npsi0, MF<T> mf, const double* w,
                      const double* wuv, void* pout, void* pmf, void* pdel, double* partial, hipStream_t st,
                      const WJob& job);

int hpsi_mfma(const hfb_ctx* c, const void* psi, int ns, int npsi0, MF<float> mf, const double* w,
              const double* wuv, void* pout, void* pmf, void* pdel, double* partial, hipStream_t st,
              const WJob& job) {
  // the weights in the x launch; the energy tail in the plane launch, which
  // runs longer
  const WJob xj = weights_part(job);
  const int rows = ns + job_rows(xj);
  hipLaunchKernelGGL(hpsi_mfma_x_kernel, dim3(c->ny * (c->nz / 16) / 4, rows), dim3(256), 0, st,
                     (const cx<float>*)psi, (cx<float>*)pmf, mf, npsi0, c->ny, c->nz, float(c->dx), c->der1f[0],
                     xj);
  int rc = launch_status();
  if (rc != HFB_OK) return rc;
  return launch_hpsi_plane<float>(c, psi, ns, npsi0, mf, w, wuv, pout, pmf, pdel, partial, st, tail_part(job));
}

template <typename T>
int hpsi_all(const hfb_ctx* c, const void* psi, int ns, int npsi0, MF<T> mf, const double* w, const double* wuv,
             void* pout, void* pmf, void* pdel, double* partial, hipStream_t st, const WJob& job) {
  if constexpr (sizeof(T) == 4) {
    if (hpsi_mfma_path(c, HFB_F32))
      return hpsi_mfma(c, psi, ns, npsi0, mf, w, wuv, pout, pmf, pdel, partial, st, job);
  }
  // with a plane launch the energy tail runs there, the weights in the x launch
  const bool plane = hpsi_plane_path(c, sizeof(T) == 4 ? HFB_F32 : HFB_F64);
  int rc = launch_hpsi_axis<T, 0>(c, psi, ns, npsi0, mf, w, wuv, pout, pmf, pdel, partial, st,
                                  plane ? weights_part(job) : job);
  if (plane) {
    if (rc == HFB_OK)
      rc = launch_hpsi_plane<T>(c, psi, ns, npsi0, mf, w, wuv, pout, pmf, pdel, partial, st, tail_part(job));
    return rc;
  }
  if (rc == HFB_OK) rc = launch_hpsi_axis<T, 1>(c, psi, ns, npsi0, mf, w, wuv, pout, pmf, pdel, partial, st);
  if (rc == HFB_OK) rc = launch_hpsi_axis<T, 2>(c, psi, ns, npsi0, mf, w, wuv, pout, pmf, pdel, partial, st);
  return rc;
}

}  // namespace

extern "C" int hfb_cderv(hfb_ctx* c, int prec, int axis, int kind, const void* in, const void* bmass,
                         void* out1, void* out2, int nfields, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!in || !out1 || nfields <= 0 || axis < 0 || axis > 2 || kind < 0 || kind > 2) return HFB_E_ARG;
  if (kind >= 1 && !out2) return HFB_E_ARG;
  if (kind == 2 && !bmass) return HFB_E_ARG;
  if (prec == HFB_F64) return cderv_dispatch<double>(c, axis, kind, in, bmass, out1, out2, nfields, (hipStream_t)stream);
  if (prec == HFB_F32) return cderv_dispatch<float>(c, axis, kind, in, bmass, out1, out2, nfields, (hipStream_t)stream);
  return HFB_E_ARG;
}

extern "C" int hfb_hpsi_ntiles(hfb_ctx* c, int prec) {
  if (!c) return HFB_E_NOCTX;
  if (hpsi_mfma_path(c, prec) || hpsi_plane_path(c, prec)) return c->nx;  // one partial per y-z plane
  const int ppb = prec == HFB_F32 ? hpsi_pencils_per_block<float>() : hpsi_pencils_per_block<double>();
  return pencils_of_axis(c, 2) / ppb;
}

extern "C" int hfb_hpsi_ex(hfb_ctx* c, int prec, const void* psi, int nstates, int npsi0, const void* upot,
                        const void* bmass, const void* aq, const void* spot, const void* wlspot, const void* v_pair,
                        const double* w, const double* wuv, void* pout, void* pout_mf, void* pout_del,
                        double* partial, int* ntile_out, const hfb_weights_job* job,
                           void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!psi || !pout_mf || !upot || !bmass || !aq || !spot || !wlspot || !v_pair || !w || !wuv || nstates <= 0)
    return HFB_E_ARG;
  WJob jb;
  if (!wjob_of(c, job, jb)) return HFB_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (ntile_out) *ntile_out = hfb_hpsi_ntiles(c, prec);
  if (prec == HFB_F64) {
    MF<double> mf{(const double*)upot, (const double*)bmass, (const double*)aq,
                  (const double*)spot, (const double*)wlspot, (const double*)v_pair};
    return hpsi_all<double>(c, psi, nstates, npsi0, mf, w, wuv, pout, pout_mf, pout_del, partial, st, jb);
  }
  if (prec == HFB_F32) {
    MF<float> mf{(const float*)upot, (const float*)bmass, (const float*)aq,
                 (const float*)spot, (const float*)wlspot, (const float*)v_pair};
    return hpsi_all<float>(c, psi, nstates, npsi0, mf, w, wuv, pout, pout_mf, pout_del, partial, st, jb);
  }
  return HFB_E_ARG;
}

extern "C" int hfb_hpsi(hfb_ctx* c, int prec, const void* psi, int nstates, int npsi0, const void* upot,
                        const void* bmass, const void* aq, const void* spot, const void* wlspot, const void* v_pair,
                        const double* w, const double* wuv, void* pout, void* pout_mf, void* pout_del,
                        double* partial, int* ntile_out, void* stream) {
  return hfb_hpsi_ex(c, prec, psi, nstates, npsi0, upot, bmass, aq, spot, wlspot, v_pair, w, wuv, pout, pout_mf,
                     pout_del, partial, ntile_out, nullptr, stream);
}
