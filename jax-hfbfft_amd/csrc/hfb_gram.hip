// Subspace matrices and rotations of diagstep (static.py:261-577) on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// Gram: C_m = A^H B_m for up to four right-hand blocks sharing A = psi
// (overlap psi^H psi, H_old = psi^H hmfpsi, lambda_lin = psi^H hampsi,
// G_old = psi^H delpsi), K = 2*nx*ny*nz per state.  Each wavefront owns a
// 16x16 output tile and a K-chunk; chunks are summed in a fixed order by a
// second kernel, so results are deterministic.
//
// Rotation: out_j = sum_i psi_i U[i][j] for two matrices at once (psi U and
// psi (U lambda+), i.e. the rotated states and the Lagrange field) computed
// as D = U^T psi^T so that each output row is a contiguous run of points.
//
// MFMA f64 16x16x4 lane layout (checked by hfb_selftest_mfma_f64):
//   A: lane l holds A[l%16][l/16];  B: lane l holds B[l/16][l%16];
//   D: lane l, element r holds D[l/16 + 4r][l%16].
#include <math.h>

#include "hfb_ctx.h"

using namespace hfb;

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <typename T>
__device__ __forceinline__ cx<double> ld64(const cx<T>* p) {
  const cx<T> v = *p;
  return {double(v.x), double(v.y)};
}

struct GramArgs {
  const void* A;
  const void* B[4];
  int nm;
  int nst;
  int64_t K;
  int64_t kc;   // K per chunk, multiple of 16
  int np;       // nst padded to 16
  double* part; // [chunk][m][np][np] complex
};

template <typename T, int NM>
__global__ void __launch_bounds__(64) gram_kernel(GramArgs g) {
  const int chunk = blockIdx.x;
  const int ntj = g.np / 16;
  const int ti = blockIdx.y / ntj, tj = blockIdx.y % ntj;
  const int l = threadIdx.x;
  const int grp = l >> 4, c = l & 15;
  const int i = ti * 16 + c;
  const int j = tj * 16 + c;
  const int64_t q = g.kc / 4;
  const int64_t kb = (int64_t)chunk * g.kc + grp * q;
  const int64_t kend = g.K;
  const cx<T>* ar = (const cx<T>*)g.A + (int64_t)(i < g.nst ? i : 0) * g.K;
  const cx<T>* br[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) br[m] = (const cx<T>*)g.B[m] + (int64_t)(j < g.nst ? j : 0) * g.K;
  const bool iok = i < g.nst, jok = j < g.nst;
  v4d cre[NM], cim[NM];
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    cre[m] = v4d{0.0, 0.0, 0.0, 0.0};
    cim[m] = v4d{0.0, 0.0, 0.0, 0.0};
  }
  for (int64_t s = 0; s < q; s += 4) {
    cx<double> a[4], b[NM][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = kb + s + u;
      const bool kin = k < kend && (s + u) < q;
      a[u] = (iok && kin) ? ld64(ar + k) : cx<double>{0.0, 0.0};
#pragma unroll
      for (int m = 0; m < NM; ++m) b[m][u] = (jok && kin) ? ld64(br[m] + k) : cx<double>{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        // conj(a) b = (ar br + ai bi) + i (ar bi - ai br)
        cre[m] = mfma(a[u].x, b[m][u].x, cre[m]);
        cre[m] = mfma(a[u].y, b[m][u].y, cre[m]);
        cim[m] = mfma(a[u].x, b[m][u].y, cim[m]);
        cim[m] = mfma(-a[u].y, b[m][u].x, cim[m]);
      }
    }
  }
  const int np = g.np;
#pragma unroll
  for (int m = 0; m < NM; ++m) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = ti * 16 + grp + 4 * r;
      const int col = tj * 16 + c;
      double* dst = g.part + 2 * ((((int64_t)chunk * NM + m) * np + row) * np + col);
      dst[0] = cre[m][r];
      dst[1] = cim[m][r];
    }
  }
}

__global__ void gram_reduce_kernel(const double* __restrict__ part, int nchunk, int nm, int np, int nst,
                                   double scale, double* __restrict__ out) {
  // out: [m][nst][nst] complex
  const int64_t total = (int64_t)nm * nst * nst;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int m = e / ((int64_t)nst * nst);
    const int ij = e % ((int64_t)nst * nst);
    const int i = ij / nst, j = ij % nst;
    double re = 0.0, im = 0.0;
    for (int ch = 0; ch < nchunk; ++ch) {
      const double* p = part + 2 * ((((int64_t)ch * nm + m) * np + i) * np + j);
      re += p[0];
      im += p[1];
    }
    out[2 * e] = re * scale;
    out[2 * e + 1] = im * scale;
  }
}

int gram_nchunk(int np, int64_t K) {
  const int ntile = (np / 16) * (np / 16);
  int nchunk = (1024 + ntile - 1) / ntile;
  const int64_t maxch = (K + 63) / 64;
  if (nchunk > maxch) nchunk = (int)maxch;
  return nchunk < 1 ? 1 : nchunk;
}

// ---------------------------------------------------------------------------
// Rotation: out1_j = sum_i psi_i U1[i][j], out2_j = sum_i psi_i U2[i][j]
// ---------------------------------------------------------------------------
template <typename T, int NJT>
__global__ void __launch_bounds__(256) rotate_kernel(const cx<T>* __restrict__ psi, const double* __restrict__ U1,
                                                     const double* __restrict__ U2, cx<T>* __restrict__ out1,
                                                     cx<T>* __restrict__ out2, int nst, int64_t K) {
  // U (nst x nst complex, row-major [i][j]) staged in LDS as padded (np x np)
  extern __shared__ double lds[];
  const int np = (nst + 15) & ~15;
  double* u1 = lds;                       // 2*np*np
  double* u2 = lds + 2 * (int64_t)np * np;
  for (int e = threadIdx.x; e < np * np; e += blockDim.x) {
    const int i = e / np, j = e % np;
    const bool ok = i < nst && j < nst;
    u1[2 * e] = ok ? U1[2 * (i * nst + j)] : 0.0;
    u1[2 * e + 1] = ok ? U1[2 * (i * nst + j) + 1] : 0.0;
    if (U2) {
      u2[2 * e] = ok ? U2[2 * (i * nst + j)] : 0.0;
      u2[2 * e + 1] = ok ? U2[2 * (i * nst + j) + 1] : 0.0;
    }
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const int l = threadIdx.x & 63;
  const int grp = l >> 4, c = l & 15;
  const int jg0 = blockIdx.y * NJT;          // first j-tile of this block
  const int njt = min(NJT, np / 16 - jg0);
  const int64_t nblk = (K + 15) / 16;
  for (int64_t blk = (int64_t)blockIdx.x * 4 + wave; blk < nblk; blk += (int64_t)gridDim.x * 4) {
    const int64_t k = blk * 16 + c;
    v4d d1r[NJT], d1i[NJT], d2r[NJT], d2i[NJT];
#pragma unroll
    for (int t = 0; t < NJT; ++t) {
      d1r[t] = v4d{0.0, 0.0, 0.0, 0.0};
      d1i[t] = d1r[t];
      d2r[t] = d1r[t];
      d2i[t] = d1r[t];
    }
    for (int i0 = 0; i0 < np; i0 += 4) {
      const int i = i0 + grp;  // B operand: psi^T[i][k]
      cx<double> b = (i < nst && k < K) ? ld64(psi + (int64_t)i * K + k) : cx<double>{0.0, 0.0};
#pragma unroll
      for (int t = 0; t < NJT; ++t) {
        if (t < njt) {
          const int j = (jg0 + t) * 16 + c;  // A operand: U^T[j][i] = U[i][j]
          const double* pa = u1 + 2 * ((int64_t)i * np + j);
          const double ar = pa[0], ai = pa[1];
          d1r[t] = mfma(ar, b.x, d1r[t]);
          d1r[t] = mfma(-ai, b.y, d1r[t]);
          d1i[t] = mfma(ar, b.y, d1i[t]);
          d1i[t] = mfma(ai, b.x, d1i[t]);
          if (U2) {
            const double* pb = u2 + 2 * ((int64_t)i * np + j);
            const double br = pb[0], bi = pb[1];
            d2r[t] = mfma(br, b.x, d2r[t]);
            d2r[t] = mfma(-bi, b.y, d2r[t]);
            d2i[t] = mfma(br, b.y, d2i[t]);
            d2i[t] = mfma(bi, b.x, d2i[t]);
          }
        }
      }
    }
    if (k < K) {
#pragma unroll
      for (int t = 0; t < NJT; ++t) {
        if (t < njt) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int j = (jg0 + t) * 16 + grp + 4 * r;
            if (j < nst) {
              out1[(int64_t)j * K + k] = cx<T>{T(d1r[t][r]), T(d1i[t][r])};
              if (U2) out2[(int64_t)j * K + k] = cx<T>{T(d2r[t][r]), T(d2i[t][r])};
            }
          }
        }
      }
    }
  }
}

}  // namespace

extern "C" int hfb_gram_workspace(int nst, long long K) {
  const int np = (nst + 15) & ~15;
  return gram_nchunk(np, K) * 4 * np * np * 2;  // doubles for up to four matrices
}

extern "C" int hfb_gram(int prec, const void* A, const void* const* B, int nm, int nst, long long K, double scale,
                        double* part, double* out, void* stream) {
  if (!A || !B || !part || !out || nm < 1 || nm > 4 || nst <= 0 || K <= 0) return HFB_E_ARG;
  for (int m = 0; m < nm; ++m)
    if (!B[m]) return HFB_E_ARG;
  GramArgs g{};
  g.A = A;
  for (int m = 0; m < nm; ++m) g.B[m] = B[m];
  g.nm = nm;
  g.nst = nst;
  g.K = K;
  g.np = (nst + 15) & ~15;
  const int nchunk = gram_nchunk(g.np, K);
  int64_t kc = (K + nchunk - 1) / nchunk;
  kc = (kc + 15) & ~(int64_t)15;
  g.kc = kc;
  g.part = part;
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(nchunk, (g.np / 16) * (g.np / 16));
#define HFB_GR(TT, NMV) hipLaunchKernelGGL((gram_kernel<TT, NMV>), grid, dim3(64), 0, st, g)
  if (prec == HFB_F64) {
    if (nm == 1) HFB_GR(double, 1);
    else if (nm == 2) HFB_GR(double, 2);
    else if (nm == 3) HFB_GR(double, 3);
    else HFB_GR(double, 4);
  } else if (prec == HFB_F32) {
    if (nm == 1) HFB_GR(float, 1);
    else if (nm == 2) HFB_GR(float, 2);
    else if (nm == 3) HFB_GR(float, 3);
    else HFB_GR(float, 4);
  } else {
    return HFB_E_ARG;
  }
#undef HFB_GR
  int rc = launch_status();
  if (rc != HFB_OK) return rc;
  const int64_t tot = (int64_t)nm * nst * nst;
  const int nb = (int)((tot + 255) / 256);
  hipLaunchKernelGGL(gram_reduce_kernel, dim3(nb < 1 ? 1 : nb), dim3(256), 0, st, part, nchunk, nm, g.np, nst, scale,
                     out);
  return launch_status();
}

extern "C" int hfb_rotate(int prec, const void* psi, const double* U1, const double* U2, void* out1, void* out2,
                          int nst, long long K, void* stream) {
  if (!psi || !U1 || !out1 || nst <= 0 || K <= 0) return HFB_E_ARG;
  if (U2 && !out2) return HFB_E_ARG;
  const int np = (nst + 15) & ~15;
  constexpr int NJT = 4;
  const int jgroups = (np / 16 + NJT - 1) / NJT;
  const size_t lds = sizeof(double) * 2 * 2 * (size_t)np * np;
  if (lds > 160 * 1024) return HFB_E_SIZE;
  const int64_t nblk = (K + 15) / 16;
  int gx = (int)((nblk + 4 * 8 - 1) / (4 * 8));  // ~8 point blocks per wave
  if (gx < 1) gx = 1;
  hipStream_t st = (hipStream_t)stream;
  if (prec == HFB_F64) {
    if (lds > 64 * 1024)
      hipFuncSetAttribute((const void*)rotate_kernel<double, NJT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((rotate_kernel<double, NJT>), dim3(gx, jgroups), dim3(256), lds, st, (const cx<double>*)psi, U1,
                       U2, (cx<double>*)out1, (cx<double>*)out2, nst, (int64_t)K);
  } else if (prec == HFB_F32) {
    if (lds > 64 * 1024)
      hipFuncSetAttribute((const void*)rotate_kernel<float, NJT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((rotate_kernel<float, NJT>), dim3(gx, jgroups), dim3(256), lds, st, (const cx<float>*)psi, U1,
                       U2, (cx<float>*)out1, (cx<float>*)out2, nst, (int64_t)K);
  } else {
    return HFB_E_ARG;
  }
  return launch_status();
}
