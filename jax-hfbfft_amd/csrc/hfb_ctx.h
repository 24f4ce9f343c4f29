// Library context: grid geometry and the device tables every kernel shares.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hfb_common.h"
#include "../../include/hfb_hip.h"

struct hfb_ctx {
  int dev;
  int nx, ny, nz;
  double dx, dy, dz;
  int periodic;
  int nx2, ny2, nz2;  // Coulomb grid
  // twiddles exp(-2 pi i m / N) for each distinct length used
  int tw_len[6];
  hfb::cx<double>* tw64[6];
  hfb::cx<float>* tw32[6];
  int ntw;
  double* der1[3];  // dense first/second derivative matrices (n x n), FP64
  double* der2[3];
  float* der1f[3];  // der1 in FP32 (the matrix-core derivatives of the FP32 spinor kernels)
  hfb::cx<double>* qhat;  // FFT of the Coulomb kernel on the doubled grid
  // arrival counters of the single-pass reductions (zero between launches;
  // regions below), so calls on one context must not overlap in time
  unsigned int* sync;
};

// Counter slots of hfb_ctx::sync.  A device-scope fence writes back and
// invalidates an XCD's L2, so arrival counting only pays off for launches
// of a few workgroups (it costs more than a separate launch for wide ones).
enum : int {
  HFB_SYNC_DIAG = 0,      // diag_kernel: one counter for its two isospin blocks
  HFB_SYNC_DIAG_U = 1,    // diag_kernel split calls: U_rho handoff flag per isospin (two slots)
  HFB_SYNC_DIAG_CTL = 3,  // wide Loewdin path: per isospin the series degree or 0 = one-workgroup kernel (two slots)
  HFB_SYNC_COUNT = 64
};

namespace hfb {

inline int hip_status(hipError_t e) { return e == hipSuccess ? HFB_OK : HFB_E_HIP; }

template <typename T>
const cx<T>* twiddles(const hfb_ctx* c, int n);
template <>
inline const cx<double>* twiddles<double>(const hfb_ctx* c, int n) {
  for (int i = 0; i < c->ntw; ++i)
    if (c->tw_len[i] == n) return c->tw64[i];
  return nullptr;
}
template <>
inline const cx<float>* twiddles<float>(const hfb_ctx* c, int n) {
  for (int i = 0; i < c->ntw; ++i)
    if (c->tw_len[i] == n) return c->tw32[i];
  return nullptr;
}

inline int axis_len(const hfb_ctx* c, int axis) { return axis == 0 ? c->nx : (axis == 1 ? c->ny : c->nz); }
inline double axis_d(const hfb_ctx* c, int axis) { return axis == 0 ? c->dx : (axis == 1 ? c->dy : c->dz); }

// last launch status, mapped to the library's codes
inline int launch_status() { return hip_status(hipGetLastError()); }

// Allow a kernel the 160 KB of LDS less its static LDS as dynamic shared
// memory; the attribute is set once per kernel (outside any stream capture
// in practice, since the first call of every kernel is an eager one).
inline void allow_full_lds(const void* kern) {
  static const void* done[128];
  static int ndone = 0;
  for (int i = 0; i < ndone; ++i)
    if (done[i] == kern) return;
  hipFuncAttributes fa{};
  size_t stat = 8 * 1024;
  if (hipFuncGetAttributes(&fa, kern) == hipSuccess) stat = fa.sharedSizeBytes;
  hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(160 * 1024 - stat));
  if (ndone < 128) done[ndone++] = kern;
}

}  // namespace hfb

// Dispatch a functor over the supported pencil lengths.
#define HFB_DISPATCH_N(n, ...)                    \
  switch (n) {                                   \
    case 24: { constexpr int N_ = 24; __VA_ARGS__; break; } \
    case 32: { constexpr int N_ = 32; __VA_ARGS__; break; } \
    case 48: { constexpr int N_ = 48; __VA_ARGS__; break; } \
    default: return HFB_E_SIZE;                  \
  }

#define HFB_DISPATCH_N2(n, ...)                   \
  switch (n) {                                   \
    case 48: { constexpr int N_ = 48; __VA_ARGS__; break; } \
    case 64: { constexpr int N_ = 64; __VA_ARGS__; break; } \
    case 96: { constexpr int N_ = 96; __VA_ARGS__; break; } \
    default: return HFB_E_SIZE;                  \
  }
