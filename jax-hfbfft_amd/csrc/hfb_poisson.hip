// Coulomb potential by FFT Green's function on the doubled grid
// (coulomb.py:14-75).  A 3-D transform is three pencil passes; the
// Poisson solve fuses zero-padding into the first pass, the kernel product
// into the x pass (forward, multiply, inverse) and the corner extraction
// into the last pass.  On a cubic doubled grid the solve takes three
// launches that skip the known zeros: y-z planes of the nx nonzero x
// planes (forward along z then y, the plane in LDS), x pencils (forward
// with the upper half implicitly zero, kernel product, inverse, lower half
// kept), y-z planes again (inverse along y then z, corner extracted).
#include <math.h>

#include <mutex>

#include "hfb_ctx.h"
#include "hfb_poisson.h"

using namespace hfb;

namespace {

enum PassMode {
  PM_QGEN = 0,     // z pass: generate the real 1/r kernel and transform it
  PM_PAD = 1,      // z pass: zero-padded proton density in, forward
  PM_FWD = 2,      // forward in place
  PM_MUL = 3,      // x pass: forward, multiply by the kernel, inverse
  PM_INV = 4,      // inverse in place
  PM_EXTRACT = 5   // z pass: inverse, real part of the corner / N2 out
};

struct PoisArgs {
  int n2x, n2y, n2z;   // doubled (or periodic) grid
  int nx, ny, nz;      // physical grid
  double dx, dy, dz;
  int periodic;
  double scale;        // PM_MUL: e2*wxyz (non-periodic) or 4 pi e2 (periodic); PM_EXTRACT: 1/N2
  const double* rho;   // PM_PAD: proton density (nx,ny,nz)
  double* wcoul;       // PM_EXTRACT output (nx,ny,nz)
  const cx<double>* qhat;
};

__device__ __forceinline__ int wrap_idx(int i, int n) { return i <= n / 2 ? i : i - n; }

template <int N, int AXIS, int MODE>
__global__ void __launch_bounds__(128) pois_pass(cx<double>* __restrict__ buf, PoisArgs a,
                                                 const cx<double>* __restrict__ twg) {
  constexpr int P = 8;
  __shared__ cx<double> tw[N];
  __shared__ cx<double> b0[N * P], b1[N * P];
  const int tile = blockIdx.x;
  const Tile t = make_tile<AXIS>(tile, a.n2x, a.n2y, a.n2z, P);
  load_twiddles(tw, twg, N);
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    int k, p;
    elem_of<AXIS, N, P>(e, k, p);
    const int64_t off = t.base + k * t.sk + p * t.sp;
    cx<double> v;
    if constexpr (MODE == PM_QGEN || MODE == PM_PAD) {
      int x, y, z;
      coords_of<AXIS>(tile, k, p, a.n2x, a.n2y, a.n2z, P, x, y, z);
      if constexpr (MODE == PM_QGEN) {
        // q(r) = 1/|r| with wrapped indices, q(0) = 2.84/(dx dy dz)^(1/3) (coulomb.py:14-48)
        const double ix = a.dx * wrap_idx(x, a.n2x), iy = a.dy * wrap_idx(y, a.n2y), iz = a.dz * wrap_idx(z, a.n2z);
        const double r2 = ix * ix + iy * iy + iz * iz;
        v.x = (x == 0 && y == 0 && z == 0) ? 2.84 / cbrt(a.dx * a.dy * a.dz) : 1.0 / sqrt(r2);
        v.y = 0.0;
      } else {
        const bool in = x < a.nx && y < a.ny && z < a.nz;
        v.x = in ? a.rho[((int64_t)x * a.ny + y) * a.nz + z] : 0.0;
        v.y = 0.0;
      }
    } else {
      v = buf[off];
    }
    b0[k * P + p] = v;
  }
  __syncthreads();
  cx<double>* r;
  if constexpr (MODE == PM_INV || MODE == PM_EXTRACT) {
    r = fft_lds<double, N, P, +1>(b0, b1, tw, P);
  } else {
    r = fft_lds<double, N, P, -1>(b0, b1, tw, P);
  }
  if constexpr (MODE == PM_MUL) {
    for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
      const int k = e / P, p = e % P;
      const int64_t off = t.base + k * t.sk + p * t.sp;
      cx<double> f;
      if (a.periodic) {
        int x, y, z;
        coords_of<AXIS>(tile, k, p, a.n2x, a.n2y, a.n2z, P, x, y, z);
        const double kx = 2.0 * M_PI * wrap_idx(x, a.n2x) / (a.n2x * a.dx);
        const double ky = 2.0 * M_PI * wrap_idx(y, a.n2y) / (a.n2y * a.dy);
        const double kz = 2.0 * M_PI * wrap_idx(z, a.n2z) / (a.n2z * a.dz);
        const double k2 = kx * kx + ky * ky + kz * kz;
        f.x = (x == 0 && y == 0 && z == 0) ? 0.0 : a.scale / k2;
        f.y = 0.0;
      } else {
        f = cscale(a.qhat[off], a.scale);
      }
      r[e] = cmul(r[e], f);
    }
    __syncthreads();
    cx<double>* o = (r == b0) ? b1 : b0;
    r = fft_lds<double, N, P, +1>(r, o, tw, P);
  }
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    int k, p;
    elem_of<AXIS, N, P>(e, k, p);
    const int64_t off = t.base + k * t.sk + p * t.sp;
    if constexpr (MODE == PM_EXTRACT) {
      int x, y, z;
      coords_of<AXIS>(tile, k, p, a.n2x, a.n2y, a.n2z, P, x, y, z);
      if (x < a.nx && y < a.ny && z < a.nz) a.wcoul[((int64_t)x * a.ny + y) * a.nz + z] = r[k * P + p].x * a.scale;
    } else {
      buf[off] = r[k * P + p];
    }
  }
}

// ----- cubic doubled grid, three launches -------------------------------
template <int N>
__global__ void __launch_bounds__(N * PP_L) pois_plane_fwd(cx<double>* __restrict__ out, const double* __restrict__ rho,
                                                           int nx, int ny, int nz, const cx<double>* __restrict__ tw) {
  extern __shared__ unsigned char pp_lds[];
  pois_fwd_body<N>(out, rho, nx, ny, nz, tw, blockIdx.x, (cx<double>*)pp_lds);
}

// PP_PW pencils per 64-lane wavefront
template <int N>
__global__ void __launch_bounds__(64) pois_pencil_mul(cx<double>* __restrict__ buf, const cx<double>* __restrict__ qhat,
                                                      double scale, int nx, const cx<double>* __restrict__ tw) {
  pois_mul_body<N>(buf, qhat, scale, nx, tw, blockIdx.x * PP_PW + threadIdx.x / PP_L, threadIdx.x % PP_L);
}

template <int N>
__global__ void __launch_bounds__(N * PP_L) pois_plane_inv(const cx<double>* __restrict__ in, double* __restrict__ wcoul,
                                                           double scale, int ny, int nz,
                                                           const cx<double>* __restrict__ tw) {
  extern __shared__ unsigned char pp_lds[];
  pois_inv_body<N>(in, wcoul, scale, ny, nz, tw, blockIdx.x, (cx<double>*)pp_lds);
}

int poisson_planes(const hfb_ctx* c, const double* rho_p, double* wcoul, double e2, cx<double>* buf,
                   hipStream_t st) {
  const int n2 = c->nx2;
  const cx<double>* tw = twiddles<double>(c, n2);
  if (!tw) return HFB_E_SIZE;
  const double sc = e2 * c->dx * c->dy * c->dz;
  const double inv = 1.0 / ((double)n2 * n2 * n2);
#define HFB_PPL(NN)                                                                                         \
  case NN: {                                                                                                \
    const size_t lds = sizeof(cx<double>) * (size_t)NN * (NN + 1);                                          \
    allow_full_lds((const void*)pois_plane_fwd<NN>);                                                        \
    allow_full_lds((const void*)pois_plane_inv<NN>);                                                        \
    hipLaunchKernelGGL(pois_plane_fwd<NN>, dim3(c->nx), dim3(NN * PP_L), lds, st, buf, rho_p, c->nx, c->ny,   \
                       c->nz, tw);                                                                          \
    int rc = launch_status();                                                                               \
    if (rc != HFB_OK) return rc;                                                                            \
    hipLaunchKernelGGL(pois_pencil_mul<NN>, dim3(NN * NN / PP_PW), dim3(64), 0, st, buf, c->qhat, sc, c->nx,  \
                       tw);                                                                                 \
    rc = launch_status();                                                                                   \
    if (rc != HFB_OK) return rc;                                                                            \
    hipLaunchKernelGGL(pois_plane_inv<NN>, dim3(c->nx), dim3(NN * PP_L), lds, st, buf, wcoul, inv, c->ny,     \
                       c->nz, tw);                                                                          \
    return launch_status();                                                                                 \
  }
  switch (n2) {
    HFB_PPL(48)
    HFB_PPL(64)
    HFB_PPL(96)
    default:
      return HFB_E_SIZE;
  }
#undef HFB_PPL
}

bool planes_apply(const hfb_ctx* c) { return pois_planes_apply(c); }

template <int AXIS, int MODE>
int launch_pass(const hfb_ctx* c, cx<double>* buf, const PoisArgs& a, hipStream_t st) {
  constexpr int P = 8;
  const int n = AXIS == 0 ? a.n2x : (AXIS == 1 ? a.n2y : a.n2z);
  const int nt = tiles_per_field<AXIS>(a.n2x, a.n2y, a.n2z, P);
  const cx<double>* tw = twiddles<double>(c, n);
  if (!tw) return HFB_E_SIZE;
  switch (n) {
#define HFB_PP(NN)                                                                                   \
  case NN:                                                                                           \
    hipLaunchKernelGGL((pois_pass<NN, AXIS, MODE>), dim3(nt), dim3(128), 0, st, buf, a, tw); \
    break;
    HFB_PP(24)
    HFB_PP(32)
    HFB_PP(48)
    HFB_PP(64)
    HFB_PP(96)
#undef HFB_PP
    default:
      return HFB_E_SIZE;
  }
  return launch_status();
}

PoisArgs base_args(const hfb_ctx* c) {
  PoisArgs a{};
  a.n2x = c->nx2;
  a.n2y = c->ny2;
  a.n2z = c->nz2;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.dx = c->dx;
  a.dy = c->dy;
  a.dz = c->dz;
  a.periodic = c->periodic;
  a.qhat = c->qhat;
  return a;
}

}  // namespace

// FFT of the non-periodic Coulomb kernel, once per context (coulomb.py:23-52).
int hfb_coulomb_setup(hfb_ctx* c) {
  if (c->periodic) return HFB_OK;
  const size_t n2 = (size_t)c->nx2 * c->ny2 * c->nz2;
  if (hipMalloc(&c->qhat, sizeof(cx<double>) * n2) != hipSuccess) return HFB_E_HIP;
  PoisArgs a = base_args(c);
  int rc = launch_pass<2, PM_QGEN>(c, c->qhat, a, 0);
  if (rc == HFB_OK) rc = launch_pass<1, PM_FWD>(c, c->qhat, a, 0);
  if (rc == HFB_OK) rc = launch_pass<0, PM_FWD>(c, c->qhat, a, 0);
  if (rc == HFB_OK) rc = hip_status(hipDeviceSynchronize());
  return rc;
}

extern "C" int hfb_poisson(hfb_ctx* c, const double* rho, double* wcoul, double e2, void* work, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!rho || !wcoul || !work) return HFB_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  cx<double>* buf = (cx<double>*)work;
  const int64_t nvox = (int64_t)c->nx * c->ny * c->nz;
  if (planes_apply(c)) return poisson_planes(c, rho + nvox, wcoul, e2, buf, st);
  PoisArgs a = base_args(c);
  a.rho = rho + nvox;  // proton density rho[1] (coulomb.py:66)
  a.wcoul = wcoul;
  int rc = launch_pass<2, PM_PAD>(c, buf, a, st);
  if (rc == HFB_OK) rc = launch_pass<1, PM_FWD>(c, buf, a, st);
  a.scale = c->periodic ? 4.0 * M_PI * e2 : e2 * c->dx * c->dy * c->dz;
  if (rc == HFB_OK) rc = launch_pass<0, PM_MUL>(c, buf, a, st);
  if (rc == HFB_OK) rc = launch_pass<1, PM_INV>(c, buf, a, st);
  a.scale = 1.0 / ((double)c->nx2 * c->ny2 * c->nz2);
  if (rc == HFB_OK) rc = launch_pass<2, PM_EXTRACT>(c, buf, a, st);
  return rc;
}

// ---------------------------------------------------------------------------
// No-allocation variant: the doubled-grid work array lives in a file-scope
// device array (one copy per GPU, held until the process ends) instead of a
// caller workspace.  In host code the array's name is not a device address:
// its address on each device comes from hipGetSymbolAddress, taken once per
// device when a context is created there (hfb_ctx_create), and only that
// address is ever passed to kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[2 * HFB_POISSON_SCRATCH_MAX];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Called by hfb_ctx_create with `device` current.
int hfb_poisson_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return HFB_E_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return HFB_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HFB_E_HIP;
  g_scratch_ptr[device] = p;
  return HFB_OK;
}

extern "C" int hfb_poisson_noalloc(hfb_ctx* c, const double* rho, double* wcoul, double e2, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!rho || !wcoul) return HFB_E_ARG;
  if ((int64_t)c->nx2 * c->ny2 * c->nz2 > HFB_POISSON_SCRATCH_MAX) return HFB_E_SIZE;
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (c->dev >= 0 && c->dev < kMaxDevices) p = g_scratch_ptr[c->dev];
  }
  if (!p) return HFB_E_NOCTX;
  // every element of the work array that a pass reads is written earlier in
  // the same call, so nothing is cleared here
  return hfb_poisson(c, rho, wcoul, e2, p, stream);
}
