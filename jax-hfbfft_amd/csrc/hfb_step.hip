// Damped gradient step with the FFT-diagonal preconditioner
// (static.py:93-146 e0dmp_gt_zero, levels.py:397-437 laplace).
//
// The 3-D transform of the residual takes three passes over each spin field
// when ny == nz: a y-z plane pass (residual r formed from ps1 and lagrange
// or h_exp*psi, forward transforms along z then y with the plane held in
// LDS between them), an x pencil pass (forward, divide by the
// preconditioner, inverse), and a y-z plane pass (inverse along y then z,
// psi <- psi - x0dmp * P r).  Other grids take five pencil passes: z (in),
// y (forward), x (divide), y (inverse), z (out).  A scratch array of the
// spinor block's size carries the transform between passes; along an axis
// that has been transformed, element k of a pencil holds frequency k.  Each
// pencil sits in a group of L lanes (LaneFFT layout): a forward transform
// leaves element k1 of lane j at frequency position k1 + N1 bitrev(j), an
// inverse one reads from there.
#include <math.h>

#include "hfb_ctx.h"
#include "hfb_jobs.h"

using namespace hfb;

namespace {

enum StepMode { SM_IN = 0, SM_FWD = 1, SM_DIV = 2, SM_INV = 3, SM_OUT = 4 };

struct StepArgs {
  int nx, ny, nz;
  const void* ps1;       // w h psi - wuv v_pair psi
  const void* lag;       // Lagrange field
  void* psi;             // in/out
  void* work;            // transform scratch, same shape as psi
  const double* hsum;    // (S,2): (sum Re psi* h psi, sum Re psi* ps1), or (S, hsum_nt, 2) partials
  int hsum_nt;           // 0: hsum holds totals
  double wxyz;
  int use_lag;           // (iteration > 1) && !tbcs  (static.py:99-104)
  const double* pw;      // max(weight, 0.1) per state
  const double* pshift;  // 0.5*max(weightuv, 0.1)*max(v_pair) per state
  double e0, h2ma, x0;
  double kfx, kfy, kfz;  // 2 pi / (d n) per axis
  double* fpart;         // (2S fields, ntile, 2) partials of |r|^2, |lag - h_exp psi|^2
};

// Total j of state s from per-state sums (nt == 0) or from nt tile
// partials per state, added in tile order.
__device__ __forceinline__ double state_sum(const double* __restrict__ p, int nt, int s, int j) {
  if (nt <= 0) return p[2 * s + j];
  const double* q = p + (int64_t)s * nt * 2 + j;
  double v = 0.0;
  for (int t = 0; t < nt; ++t) v += q[2 * t];
  return v;
}

constexpr int STEP_L = 4;                 // lanes per pencil
constexpr int STEP_PPB = 4 * (64 / STEP_L);  // pencils per block

inline int step_pencils(const hfb_ctx* c, int axis) {
  return axis == 0 ? c->ny * c->nz : (axis == 1 ? c->nx * c->nz : c->nx * c->ny);
}

template <typename T, int N, int AXIS, int MODE>
__global__ void __launch_bounds__(256) step_lane(StepArgs a, const cx<T>* __restrict__ tw) {
  constexpr int L = STEP_L, N1 = N / L, PW = 64 / L;
  __shared__ double red[4 * 2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j = lane % L, q = lane / L;
  const int g = (blockIdx.x * 4 + wave) * PW + q;  // pencil
  const int f = blockIdx.y;                         // field = 2*state + spin
  const int s = f >> 1;
  const int64_t nvox = (int64_t)a.nx * a.ny * a.nz;
  const int64_t fo = (int64_t)f * nvox;
  int64_t base, stride;
  int c0, c1;
  if constexpr (AXIS == 0) {
    base = g;
    stride = (int64_t)a.ny * a.nz;
    c0 = g / a.nz;
    c1 = g % a.nz;
  } else if constexpr (AXIS == 1) {
    base = (int64_t)(g / a.nz) * a.ny * a.nz + g % a.nz;
    stride = a.nz;
    c0 = g / a.nz;
    c1 = g % a.nz;
  } else {
    base = (int64_t)g * a.nz;
    stride = 1;
    c0 = g / a.ny;
    c1 = g % a.ny;
  }
  LaneFFT<T, N, L> F;
  F.init(j, tw);
  cx<T>* psi = (cx<T>*)a.psi;
  cx<T>* R = (cx<T>*)a.work;
  constexpr bool IN_SPECTRAL = MODE == SM_INV || MODE == SM_OUT;  // input pencil holds frequencies
  cx<T> v[N1];
  double acc[2] = {0.0, 0.0};
  // h_exp only enters the residual without the Lagrange field and the
  // fluctuation sums, so the partials are not read otherwise
  T hexp = T(0);
  if constexpr (MODE == SM_IN)
    if (!a.use_lag || a.fpart) hexp = T(a.wxyz * state_sum(a.hsum, a.hsum_nt, s, 1));
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const int kk = IN_SPECTRAL ? F.freq(m) : j + L * m;
    const int64_t off = fo + base + (int64_t)kk * stride;
    if constexpr (MODE == SM_IN) {
      const cx<T> p1 = ((const cx<T>*)a.ps1)[off];
      const cx<T> ps = psi[off];
      const cx<T> lg = ((const cx<T>*)a.lag)[off];
      const cx<T> hp = cscale(ps, hexp);
      v[m] = a.use_lag ? csub(p1, lg) : csub(p1, hp);
      const cx<T> e2 = csub(lg, hp);
      acc[0] += double(v[m].x) * double(v[m].x) + double(v[m].y) * double(v[m].y);
      acc[1] += double(e2.x) * double(e2.x) + double(e2.y) * double(e2.y);
    } else {
      v[m] = R[off];
    }
  }
  if constexpr (MODE == SM_IN) {
    if (a.fpart) {
      block_sum<2>(acc, red);
      if (threadIdx.x == 0) {
        a.fpart[((int64_t)f * gridDim.x + blockIdx.x) * 2 + 0] = acc[0];
        a.fpart[((int64_t)f * gridDim.x + blockIdx.x) * 2 + 1] = acc[1];
      }
    }
  }
  if constexpr (IN_SPECTRAL)
    F.inv(v, tw);
  else
    F.fwd(v, tw);
  if constexpr (MODE == SM_DIV) {  // levels.py:407-434 (k^2 with the Nyquist mode kept)
    const double ntot = double(nvox);
    const int my = AXIS == 0 ? (c0 <= a.ny / 2 ? c0 : a.ny - c0) : 0;
    const int mz = AXIS == 0 ? (c1 <= a.nz / 2 ? c1 : a.nz - c1) : 0;
    const double kyz = (my * a.kfy) * (my * a.kfy) + (mz * a.kfz) * (mz * a.kfz);
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int kx = F.freq(m);
      const int mx = kx <= N / 2 ? kx : N - kx;
      const double k2 = (mx * a.kfx) * (mx * a.kfx) + kyz;
      const double den = (a.pw[s] * (a.e0 + a.h2ma * k2) + a.pshift[s]) * ntot;
      v[m] = cscale(v[m], T(1.0 / den));
    }
    F.inv(v, tw);
  }
  // output: spectral positions after a forward pass, spatial after an inverse
  constexpr bool OUT_SPECTRAL = MODE == SM_IN || MODE == SM_FWD;
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const int kk = OUT_SPECTRAL ? F.freq(m) : j + L * m;
    const int64_t off = fo + base + (int64_t)kk * stride;
    if constexpr (MODE == SM_OUT)
      psi[off] = csub(psi[off], cscale(v[m], T(a.x0)));
    else
      R[off] = v[m];
  }
}

// One y-z plane (x fixed) of one spin field per workgroup, N pencils of L =
// 8 lanes.  FWD: residual along z pencils (as SM_IN), forward z transform,
// the plane through LDS, forward y transform, work[x][ky][kz].  INV: the
// reverse, ending with the update of psi.
constexpr int PLANE_L = 8;

template <typename T, int N, bool FWD>
__global__ void __launch_bounds__(N * PLANE_L) step_plane(StepArgs a, const cx<T>* __restrict__ tw) {
  constexpr int L = PLANE_L, N1 = N / L, NT = N * L, P = N + 1;  // P: LDS row pitch
  __shared__ cx<T> pl[N * P];
  __shared__ double red[(NT / 64) * 2];
  const int j = threadIdx.x % L, q = threadIdx.x / L;  // lane in the group, pencil
  const int x = blockIdx.x, f = blockIdx.y, s = f >> 1;
  const int64_t nvox = (int64_t)a.nx * N * N;
  const int64_t pb = (int64_t)f * nvox + (int64_t)x * N * N;  // plane base
  LaneFFT<T, N, L> F;
  F.init(j, tw);
  cx<T>* psi = (cx<T>*)a.psi;
  cx<T>* R = (cx<T>*)a.work;
  cx<T> v[N1];
  if constexpr (FWD) {
    double acc[2] = {0.0, 0.0};
    // h_exp only enters the residual without the Lagrange field and the
    // fluctuation sums, so the partials are not read otherwise
    const T hexp = (!a.use_lag || a.fpart) ? T(a.wxyz * state_sum(a.hsum, a.hsum_nt, s, 1)) : T(0);
#pragma unroll
    for (int m = 0; m < N1; ++m) {  // z pencil q = y
      const int64_t off = pb + (int64_t)q * N + j + L * m;
      const cx<T> p1 = ((const cx<T>*)a.ps1)[off];
      const cx<T> ps = psi[off];
      const cx<T> lg = ((const cx<T>*)a.lag)[off];
      const cx<T> hp = cscale(ps, hexp);
      v[m] = a.use_lag ? csub(p1, lg) : csub(p1, hp);
      const cx<T> e2 = csub(lg, hp);
      acc[0] += double(v[m].x) * double(v[m].x) + double(v[m].y) * double(v[m].y);
      acc[1] += double(e2.x) * double(e2.x) + double(e2.y) * double(e2.y);
    }
    if (a.fpart) {
      block_sum<2>(acc, red);
      if (threadIdx.x == 0) {
        a.fpart[((int64_t)f * gridDim.x + x) * 2 + 0] = acc[0];
        a.fpart[((int64_t)f * gridDim.x + x) * 2 + 1] = acc[1];
      }
    }
    F.fwd(v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) pl[q * P + F.freq(m)] = v[m];  // [y][kz]
    __syncthreads();
#pragma unroll
    for (int m = 0; m < N1; ++m) v[m] = pl[(j + L * m) * P + q];  // y pencil q = kz
    F.fwd(v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) R[pb + (int64_t)F.freq(m) * N + q] = v[m];  // work[x][ky][kz]
  } else {
#pragma unroll
    for (int m = 0; m < N1; ++m) v[m] = R[pb + (int64_t)F.freq(m) * N + q];  // y pencil q = kz
    F.inv(v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) pl[(j + L * m) * P + q] = v[m];  // [y][kz]
    __syncthreads();
#pragma unroll
    for (int m = 0; m < N1; ++m) v[m] = pl[q * P + F.freq(m)];  // z pencil q = y
    F.inv(v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int64_t off = pb + (int64_t)q * N + j + L * m;
      psi[off] = csub(psi[off], cscale(v[m], T(a.x0)));
    }
  }
}

template <typename T, bool FWD>
int launch_plane(const hfb_ctx* c, const StepArgs& a, int nfields, hipStream_t st) {
  HFB_DISPATCH_N(c->ny, {
    hipLaunchKernelGGL((step_plane<T, N_, FWD>), dim3(c->nx, nfields), dim3(N_ * PLANE_L), 0, st, a,
                       twiddles<T>(c, N_));
  })
  return launch_status();
}

inline bool plane_path(const hfb_ctx* c) { return c->ny == c->nz; }

template <typename T, int AXIS, int MODE>
int launch_step(const hfb_ctx* c, const StepArgs& a, int nfields, hipStream_t st) {
  const int n = axis_len(c, AXIS);
  const int np = step_pencils(c, AXIS);
  if (np % STEP_PPB != 0) return HFB_E_SIZE;
  HFB_DISPATCH_N(n, {
    hipLaunchKernelGGL((step_lane<T, N_, AXIS, MODE>), dim3(np / STEP_PPB, nfields), dim3(256), 0, st, a,
                       twiddles<T>(c, N_));
  })
  return launch_status();
}

template <typename T>
int step_all(const hfb_ctx* c, const StepArgs& a, int ns, hipStream_t st) {
  if (plane_path(c)) {
    int rc = launch_plane<T, true>(c, a, 2 * ns, st);
    if (rc == HFB_OK) rc = launch_step<T, 0, SM_DIV>(c, a, 2 * ns, st);
    if (rc == HFB_OK) rc = launch_plane<T, false>(c, a, 2 * ns, st);
    return rc;
  }
  int rc = launch_step<T, 2, SM_IN>(c, a, 2 * ns, st);
  if (rc == HFB_OK) rc = launch_step<T, 1, SM_FWD>(c, a, 2 * ns, st);
  if (rc == HFB_OK) rc = launch_step<T, 0, SM_DIV>(c, a, 2 * ns, st);
  if (rc == HFB_OK) rc = launch_step<T, 1, SM_INV>(c, a, 2 * ns, st);
  if (rc == HFB_OK) rc = launch_step<T, 2, SM_OUT>(c, a, 2 * ns, st);
  return rc;
}

__global__ void grstep_weights_kernel(hfb_weights_job job) { grstep_weights_body(job); }

__global__ void grstep_finish_kernel(hfb_finish_job job) { grstep_finish_body(job); }

}  // namespace

extern "C" int hfb_step_ntiles(hfb_ctx* c, int prec) {
  if (!c) return HFB_E_NOCTX;
  (void)prec;
  return plane_path(c) ? c->nx : step_pencils(c, 2) / STEP_PPB;
}

extern "C" int hfb_grstep_weights(int nstates, int npsi0, int tbcs, const double* wocc, const double* wstates,
                                  const double* wguv, const double* pairwg, const double* v_pair, long long nvox,
                                  const double* vpair_max, double* w, double* wuv, double* pw, double* pshift,
                                  void* stream) {
  const hfb_weights_job job{nstates, npsi0, tbcs, wocc, wstates, wguv, pairwg, v_pair, nvox, vpair_max, w, wuv, pw,
                            pshift};
  if (!weights_job_ok(job)) return HFB_E_ARG;
  hipLaunchKernelGGL(grstep_weights_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, job);
  return launch_status();
}

extern "C" int hfb_precond_step(hfb_ctx* c, int prec, void* psi, const void* ps1, const void* lagrange,
                                int nstates, const double* hsum, int hsum_ntile, double wxyz, int use_lagrange, const double* pw,
                                const double* pshift, double e0dmp, double h2ma, double x0dmp, double* fpart,
                                void* work, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!psi || !ps1 || !lagrange || !hsum || !pw || !pshift || !work || nstates <= 0 || hsum_ntile < 0)
    return HFB_E_ARG;
  StepArgs a{};
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.ps1 = ps1;
  a.lag = lagrange;
  a.psi = psi;
  a.work = work;
  a.hsum = hsum;
  a.hsum_nt = hsum_ntile;
  a.wxyz = wxyz;
  a.use_lag = use_lagrange;
  a.pw = pw;
  a.pshift = pshift;
  a.e0 = e0dmp;
  a.h2ma = h2ma;
  a.x0 = x0dmp;
  a.kfx = 2.0 * M_PI / (c->dx * c->nx);
  a.kfy = 2.0 * M_PI / (c->dy * c->ny);
  a.kfz = 2.0 * M_PI / (c->dz * c->nz);
  a.fpart = fpart;
  hipStream_t st = (hipStream_t)stream;
  if (prec == HFB_F64) return step_all<double>(c, a, nstates, st);
  if (prec == HFB_F32) return step_all<float>(c, a, nstates, st);
  return HFB_E_ARG;
}

extern "C" int hfb_grstep_finish(int nstates, const double* hsum, int hsum_ntile, double wxyz, double* sp_energy,
                                 const double* wocc, const double* wstates, const double* fsum, int fsum_ntile,
                                 double* sp_efluct1, double* sp_efluct2, double* acc, void* stream) {
  const hfb_finish_job job{nstates, hsum, hsum_ntile, wxyz, sp_energy, wocc, wstates, fsum, fsum_ntile, sp_efluct1,
                           sp_efluct2, acc};
  if (!finish_job_ok(job)) return HFB_E_ARG;
  hipLaunchKernelGGL(grstep_finish_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, job);
  return launch_status();
}
