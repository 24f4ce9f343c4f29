// Context lifetime, error strings, partial-sum reduction and the MFMA
// lane-layout self-test.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "hfb_ctx.h"

using namespace hfb;

extern "C" const char* hfb_strerror(int code) {
  switch (code) {
    case HFB_OK: return "ok";
    case HFB_E_ARG: return "invalid argument";
    case HFB_E_SIZE: return "grid length has no compiled kernel (supported: 24, 32, 48)";
    case HFB_E_HIP: return "HIP runtime error";
    case HFB_E_NOCTX: return "null context";
    default: return "unknown error";
  }
}

namespace {

// Dense Fourier derivative matrices, grids.py:36-61 (sder, sder2).
void sder_host(int n, double d, std::vector<double>& m1, std::vector<double>& m2) {
  const int icn = (n + 1) / 2;
  const double afac = M_PI / icn;
  m1.assign((size_t)n * n, 0.0);
  m2.assign((size_t)n * n, 0.0);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      const double g = double(b - a);
      double s1 = 0.0, s2 = 0.0;
      for (int j = 1; j < icn; ++j) {
        s1 += -j * sin(j * afac * g);
        s2 += double(j) * j * cos(j * afac * g);
      }
      m1[(size_t)a * n + b] = -afac * (s1 - 0.5 * icn * sin(icn * afac * g)) / (icn * d);
      m2[(size_t)a * n + b] = -(afac * afac) * (s2 + 0.5 * icn * icn * cos(icn * afac * g)) / (icn * d * d);
    }
}

int add_twiddles(hfb_ctx* c, int n) {
  for (int i = 0; i < c->ntw; ++i)
    if (c->tw_len[i] == n) return HFB_OK;
  if (c->ntw >= 6) return HFB_E_ARG;
  std::vector<cx<double>> t64(n);
  std::vector<cx<float>> t32(n);
  for (int m = 0; m < n; ++m) {
    // exp(-2 pi i m/n) with exact values on the axes
    const double ang = -2.0 * M_PI * double(m) / double(n);
    double cr = cos(ang), ci = sin(ang);
    if ((4 * m) % n == 0) {
      const int q = (4 * m) / n;
      cr = (q == 0) ? 1.0 : (q == 2 ? -1.0 : 0.0);
      ci = (q == 1) ? -1.0 : (q == 3 ? 1.0 : 0.0);
    }
    t64[m] = {cr, ci};
    t32[m] = {float(cr), float(ci)};
  }
  const int i = c->ntw;
  if (hipMalloc(&c->tw64[i], sizeof(cx<double>) * n) != hipSuccess) return HFB_E_HIP;
  if (hipMalloc(&c->tw32[i], sizeof(cx<float>) * n) != hipSuccess) return HFB_E_HIP;
  if (hipMemcpy(c->tw64[i], t64.data(), sizeof(cx<double>) * n, hipMemcpyHostToDevice) != hipSuccess)
    return HFB_E_HIP;
  if (hipMemcpy(c->tw32[i], t32.data(), sizeof(cx<float>) * n, hipMemcpyHostToDevice) != hipSuccess)
    return HFB_E_HIP;
  c->tw_len[i] = n;
  c->ntw++;
  return HFB_OK;
}

bool supported_n(int n) { return n == 24 || n == 32 || n == 48; }

}  // namespace

int hfb_coulomb_setup(hfb_ctx* c);  // hfb_poisson.hip
int hfb_poisson_scratch_setup(int device);  // hfb_poisson.hip

extern "C" int hfb_ctx_create(hfb_ctx** out, int device, int nx, int ny, int nz, double dx, double dy,
                              double dz, int periodic) {
  if (!out) return HFB_E_ARG;
  *out = nullptr;
  if (!supported_n(nx) || !supported_n(ny) || !supported_n(nz)) return HFB_E_SIZE;
  if (!(dx > 0.0 && dy > 0.0 && dz > 0.0)) return HFB_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return HFB_E_HIP;
  hfb_ctx* c = (hfb_ctx*)calloc(1, sizeof(hfb_ctx));
  if (!c) return HFB_E_ARG;
  c->dev = device;
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->dx = dx;
  c->dy = dy;
  c->dz = dz;
  c->periodic = periodic ? 1 : 0;
  c->nx2 = periodic ? nx : 2 * nx;
  c->ny2 = periodic ? ny : 2 * ny;
  c->nz2 = periodic ? nz : 2 * nz;
  int rc = HFB_OK;
  const int lens[6] = {nx, ny, nz, c->nx2, c->ny2, c->nz2};
  for (int i = 0; i < 6 && rc == HFB_OK; ++i) rc = add_twiddles(c, lens[i]);
  const int ns[3] = {nx, ny, nz};
  const double ds[3] = {dx, dy, dz};
  for (int a = 0; a < 3 && rc == HFB_OK; ++a) {
    std::vector<double> m1, m2;
    sder_host(ns[a], ds[a], m1, m2);
    const size_t bytes = sizeof(double) * m1.size();
    if (hipMalloc(&c->der1[a], bytes) != hipSuccess || hipMalloc(&c->der2[a], bytes) != hipSuccess ||
        hipMemcpy(c->der1[a], m1.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->der2[a], m2.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
      rc = HFB_E_HIP;
    std::vector<float> f1(m1.begin(), m1.end());
    if (rc == HFB_OK && (hipMalloc(&c->der1f[a], sizeof(float) * f1.size()) != hipSuccess ||
                         hipMemcpy(c->der1f[a], f1.data(), sizeof(float) * f1.size(), hipMemcpyHostToDevice) !=
                             hipSuccess))
      rc = HFB_E_HIP;
  }
  if (rc == HFB_OK) rc = hfb_coulomb_setup(c);
  if (rc == HFB_OK) rc = hfb_poisson_scratch_setup(device);
  if (rc == HFB_OK) {
    const size_t bytes = sizeof(unsigned int) * HFB_SYNC_COUNT;
    if (hipMalloc(&c->sync, bytes) != hipSuccess || hipMemset(c->sync, 0, bytes) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
      rc = HFB_E_HIP;
  }
  if (rc != HFB_OK) {
    hfb_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return HFB_OK;
}

extern "C" int hfb_ctx_destroy(hfb_ctx* c) {
  if (!c) return HFB_E_NOCTX;
  for (int i = 0; i < c->ntw; ++i) {
    hipFree(c->tw64[i]);
    hipFree(c->tw32[i]);
  }
  for (int a = 0; a < 3; ++a) {
    if (c->der1[a]) hipFree(c->der1[a]);
    if (c->der2[a]) hipFree(c->der2[a]);
    if (c->der1f[a]) hipFree(c->der1f[a]);
  }
  if (c->qhat) hipFree(c->qhat);
  if (c->sync) hipFree(c->sync);
  free(c);
  return HFB_OK;
}

// ---------------------------------------------------------------------------
// Partial sums -> per-state totals (fixed order, deterministic).
// ---------------------------------------------------------------------------
__global__ void reduce_partials_kernel(const double* __restrict__ part, int ntile, int nv, double scale,
                                       double* __restrict__ out) {
  __shared__ double red[4 * 8];
  const int s = blockIdx.x;
  for (int j = 0; j < nv; ++j) {
    double v[1] = {0.0};
    for (int t = threadIdx.x; t < ntile; t += blockDim.x) v[0] += part[((size_t)s * ntile + t) * nv + j];
    block_sum<1>(v, red);
    if (threadIdx.x == 0) out[(size_t)s * nv + j] = scale * v[0];
  }
}

extern "C" int hfb_reduce_partials(const double* partial, int nstates, int ntile, int nv, double scale,
                                   double* out, void* stream) {
  if (!partial || !out || nstates <= 0 || ntile <= 0 || nv <= 0) return HFB_E_ARG;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(nstates), dim3(256), 0, (hipStream_t)stream, partial,
                     ntile, nv, scale, out);
  return launch_status();
}

// ---------------------------------------------------------------------------
// MFMA f64 16x16x4 lane layout self-test (one wavefront).
//   A operand: lane l holds a[row=l%16][k=l/16]
//   B operand: lane l holds b[k=l/16][col=l%16]
//   D:         lane l holds d[row=(l/16)+4*r][col=l%16], r=0..3
// d[0..256) is written with that layout, d[256..512) lane-major raw.
// ---------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void mfma_probe_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                  double* __restrict__ d) {
  const int l = threadIdx.x;
  const double av = a[(l % 16) * 4 + (l / 16)];
  const double bv = b[(l / 16) * 16 + (l % 16)];
  v4d acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) {
    d[((l / 16) + 4 * r) * 16 + (l % 16)] = acc[r];
    d[256 + l * 4 + r] = acc[r];
  }
}

extern "C" int hfb_selftest_mfma_f64(const double* a, const double* b, double* d, void* stream) {
  if (!a || !b || !d) return HFB_E_ARG;
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, a, b, d);
  return launch_status();
}

// ---------------------------------------------------------------------------
// MFMA f32 16x16x4 lane layout self-test (one wavefront): the A and B
// operands as above (a[row=l%16][k=l/16], b[k=l/16][col=l%16]); the four
// result registers of every lane are written lane-major, d[l*4 + r], so a
// test can check the D layout the h.psi kernels assume
// (d[row=4*(l/16)+r][col=l%16]).
// ---------------------------------------------------------------------------
__global__ void mfma_f32_probe_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                      float* __restrict__ d) {
  const int l = threadIdx.x;
  const float av = a[(l % 16) * 4 + (l / 16)];
  const float bv = b[(l / 16) * 16 + (l % 16)];
  v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) d[l * 4 + r] = acc[r];
}

extern "C" int hfb_selftest_mfma_f32(const float* a, const float* b, float* d, void* stream) {
  if (!a || !b || !d) return HFB_E_ARG;
  hipLaunchKernelGGL(mfma_f32_probe_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, a, b, d);
  return launch_status();
}
