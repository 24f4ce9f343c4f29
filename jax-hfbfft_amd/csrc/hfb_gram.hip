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

struct GramIso {
  const void* A;
  const void* B[4];
  int nst;
  int np;       // nst padded to 16
  double* part; // [m][np][np][chunk] complex
};

struct GramArgs {
  GramIso iso[2];
  int nm;       // number of right-hand blocks
  int bsame;    // bitmask: B[m] == A (no second load)
  int64_t K;
  int64_t kc;   // K per chunk, multiple of KS
  int nchunk;
};

constexpr int KS = 16;        // complex elements of K staged per step
constexpr int LDR = KS + 1;   // padded LDS row stride (rows sit on different banks)
constexpr int GRAM_NT = 256;  // 4 wavefronts

// One workgroup per (K-chunk, isospin block) computes every 16x16 tile of
// every C_m.  Each K-step stages A and the distinct B_m rows (KS contiguous
// elements per row, converted to double) in LDS, so every input element is
// read from HBM once; the loads of step t+1 are issued before the MFMAs of
// step t.  The four wavefronts split the output tiles; MFMA operands come
// from LDS.
template <typename T, int NM, int MAXT, int NE>
__global__ void __launch_bounds__(GRAM_NT) gram_kernel(GramArgs ga) {
  extern __shared__ double glds[];
  const GramIso& g = ga.iso[blockIdx.y];
  const int np = g.np;
  if (np == 0) return;
  const int ntile1 = np / 16;
  const int ntl = ntile1 * ntile1;
  cx<double>* sA = (cx<double>*)glds;                 // [np][LDR]
  cx<double>* sB = sA + (size_t)np * LDR;             // [NM][np][LDR]
  const int chunk = blockIdx.x;
  const int64_t k0 = (int64_t)chunk * ga.kc;
  const int64_t k1 = min(ga.K, k0 + ga.kc);
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int grp = l >> 4, c = l & 15;
  const int ntot = NM * ntl;
  const int bsame = ga.bsame;
  v4d cre[MAXT], cim[MAXT];
#pragma unroll
  for (int u = 0; u < MAXT; ++u) {
    cre[u] = v4d{0.0, 0.0, 0.0, 0.0};
    cim[u] = v4d{0.0, 0.0, 0.0, 0.0};
  }
  // register staging of one K-step: NE elements per thread and matrix
  cx<double> ra[NE], rb[NM][NE];
  auto fetch = [&](int64_t kb) {
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const int e = threadIdx.x + q * GRAM_NT;
      const int row = e / KS, kk = e % KS;
      const int64_t k = kb + kk;
      const bool ok = e < np * KS && row < g.nst && k < k1;
      ra[q] = ok ? ld64((const cx<T>*)g.A + (int64_t)row * ga.K + k) : cx<double>{0.0, 0.0};
#pragma unroll
      for (int m = 0; m < NM; ++m)
        if (!((bsame >> m) & 1))
          rb[m][q] = ok ? ld64((const cx<T>*)g.B[m] + (int64_t)row * ga.K + k) : cx<double>{0.0, 0.0};
    }
  };
  fetch(k0);
  for (int64_t kb = k0; kb < k1; kb += KS) {
#pragma unroll
    for (int q = 0; q < NE; ++q) {
      const int e = threadIdx.x + q * GRAM_NT;
      if (e < np * KS) {
        const int le = (e / KS) * LDR + e % KS;
        sA[le] = ra[q];
#pragma unroll
        for (int m = 0; m < NM; ++m)
          if (!((bsame >> m) & 1)) sB[(size_t)m * np * LDR + le] = rb[m][q];
      }
    }
    __syncthreads();
    if (kb + KS < k1) fetch(kb + KS);
#pragma unroll
    for (int u = 0; u < MAXT; ++u) {
      const int t = wave + 4 * u;
      if (t < ntot) {
        const int m = t / ntl, tt = t % ntl;
        const int ti = tt / ntile1, tj = tt % ntile1;
        const cx<double>* bb = ((bsame >> m) & 1) ? sA : sB + (size_t)m * np * LDR;
#pragma unroll
        for (int s4 = 0; s4 < KS; s4 += 4) {
          const cx<double> a = sA[(ti * 16 + c) * LDR + s4 + grp];
          const cx<double> b = bb[(tj * 16 + c) * LDR + s4 + grp];
          // conj(a) b = (ar br + ai bi) + i (ar bi - ai br)
          cre[u] = mfma(a.x, b.x, cre[u]);
          cre[u] = mfma(a.y, b.y, cre[u]);
          cim[u] = mfma(a.x, b.y, cim[u]);
          cim[u] = mfma(-a.y, b.x, cim[u]);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < MAXT; ++u) {
    const int t = wave + 4 * u;
    if (t < ntot) {
      const int m = t / ntl, tt = t % ntl;
      const int ti = tt / ntile1, tj = tt % ntile1;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = ti * 16 + grp + 4 * r;
        const int col = tj * 16 + c;
        double* dst = g.part + 2 * ((((int64_t)m * np + row) * np + col) * ga.nchunk + chunk);
        dst[0] = cre[u][r];
        dst[1] = cim[u][r];
      }
    }
  }
}

// One wavefront per output element: lanes stride over the chunks.
__global__ void __launch_bounds__(256) gram_reduce_kernel(const double* __restrict__ part, int nchunk, int nm, int np,
                                                          int nst, double scale, double* __restrict__ out) {
  const int64_t total = (int64_t)nm * nst * nst;
  const int lane = threadIdx.x & 63;
  const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t e = wid; e < total; e += nw) {
    const int m = e / ((int64_t)nst * nst);
    const int ij = e % ((int64_t)nst * nst);
    const int i = ij / nst, j = ij % nst;
    const double* p = part + 2 * (((int64_t)m * np + i) * np + j) * nchunk;
    double re = 0.0, im = 0.0;
    for (int ch = lane; ch < nchunk; ch += 64) {
      re += p[2 * ch];
      im += p[2 * ch + 1];
    }
    for (int o = 32; o > 0; o >>= 1) {
      re += __shfl_down(re, o, 64);
      im += __shfl_down(im, o, 64);
    }
    if (lane == 0) {
      out[2 * e] = re * scale;
      out[2 * e + 1] = im * scale;
    }
  }
}

int gram_nchunk(int np, int64_t K) {
  (void)np;
  int64_t nchunk = (K + 255) / 256;  // ~256 elements of K per workgroup
  if (nchunk > 256) nchunk = 256;
  return nchunk < 1 ? 1 : (int)nchunk;
}

// ---------------------------------------------------------------------------
// Rotation: out1_j = sum_i psi_i U1[i][j], out2_j = sum_i psi_i U2[i][j]
// ---------------------------------------------------------------------------
template <typename T, int NJT>
__global__ void __launch_bounds__(256) rotate_kernel(const cx<T>* psi, const double* __restrict__ U1,
                                                     const double* __restrict__ U2, cx<T>* out1, cx<T>* out2, int nst,
                                                     int64_t K) {
  // U (nst x nst complex, row-major [i][j]) staged in LDS as padded (np x np).
  // Each wavefront owns whole 16-point blocks of K: it reads psi_i for all i
  // there before writing out_j, so out1 may alias psi when one workgroup
  // column covers all j (gridDim.y == 1).
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
  const int jg0 = blockIdx.y * NJT;  // first j-tile of this block
  const int nis = np / 4;            // MFMA steps over i
  const int64_t nblk = (K + 15) / 16;
  const bool two = U2 != nullptr;
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
    for (int s0 = 0; s0 < nis; s0 += 4) {
      // B operand psi^T[i][k] for four i-steps, loaded together
      cx<double> bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = 4 * (s0 + u) + grp;
        bv[u] = (s0 + u < nis && i < nst && k < K) ? ld64(psi + (int64_t)i * K + k) : cx<double>{0.0, 0.0};
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = 4 * (s0 + u) + grp;
        const cx<double> b = bv[u];
#pragma unroll
        for (int t = 0; t < NJT; ++t) {
          const int j = (jg0 + t) * 16 + c;  // A operand: U^T[j][i] = U[i][j]
          const double* pa = u1 + 2 * ((int64_t)i * np + j);
          const double ar = pa[0], ai = pa[1];
          d1r[t] = mfma(ar, b.x, d1r[t]);
          d1r[t] = mfma(-ai, b.y, d1r[t]);
          d1i[t] = mfma(ar, b.y, d1i[t]);
          d1i[t] = mfma(ai, b.x, d1i[t]);
          if (two) {
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
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int j = (jg0 + t) * 16 + grp + 4 * r;
          if (j < nst) {
            out1[(int64_t)j * K + k] = cx<T>{T(d1r[t][r]), T(d1i[t][r])};
            if (two) out2[(int64_t)j * K + k] = cx<T>{T(d2r[t][r]), T(d2i[t][r])};
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Small blocks (nst <= 64) on the vector FP64 units.  On gfx950 the FP64
// matrix and vector rates are equal, and 16x16 MFMA tiles padded from 20 or
// 36 states waste 1.6-2.9x of the work, so blocks up to 64 states use FMAs:
// a wavefront owns a 4x4 tile of one C_m (complex, in registers) and a chunk
// of K, lanes stride over k (coalesced rows); the lane sums are combined by
// shuffles and the chunks by gram_reduce_kernel in a fixed order.  For
// B_m == A only tiles with ti <= tj are computed; the reduction mirrors the
// rest (C^H = C).
// ---------------------------------------------------------------------------
constexpr int GV_T = 4;      // tile edge
constexpr int GV_KPL = 32;   // k per lane per chunk
constexpr int GV_CHUNK = 64 * GV_KPL;

int gram_vec_nchunk(int64_t K) {
  int64_t nc = (K + GV_CHUNK - 1) / GV_CHUNK;
  return nc < 1 ? 1 : (nc > 1024 ? 1024 : (int)nc);
}

struct GramVecArgs {
  GramIso iso[2];
  int nm;
  int bsame;
  int64_t K;
  int64_t kc;      // K per chunk
  int nchunk;
  int ntile[2];    // tiles per isospin (over all m)
};

// tile index -> (m, ti, tj) for one isospin: matrices in order; for a
// matrix with B == A the upper-triangular tiles (ti <= tj) row by row.
__device__ __forceinline__ void gv_tile(int t, int nt, int nm, int bsame, int& m, int& ti, int& tj) {
  m = 0;
  for (; m < nm; ++m) {
    const bool sym = (bsame >> m) & 1;
    const int cnt = sym ? nt * (nt + 1) / 2 : nt * nt;
    if (t < cnt) break;
    t -= cnt;
  }
  const bool sym = (bsame >> m) & 1;
  if (!sym) {
    ti = t / nt;
    tj = t % nt;
    return;
  }
  ti = 0;
  while (t >= nt - ti) {
    t -= nt - ti;
    ++ti;
  }
  tj = ti + t;
}

template <typename T>
__global__ void __launch_bounds__(256) gram_vec_kernel(GramVecArgs ga) {
  const int iq = blockIdx.y;
  const GramIso& g = ga.iso[iq];
  const int nst = g.nst;
  if (nst == 0) return;
  const int lane = threadIdx.x & 63;
  const int item = blockIdx.x * 4 + (threadIdx.x >> 6);  // (tile, chunk)
  const int ntile = ga.ntile[iq];
  if (item >= ntile * ga.nchunk) return;
  const int tile = item / ga.nchunk, chunk = item % ga.nchunk;
  const int nt = (nst + GV_T - 1) / GV_T;
  int m, ti, tj;
  gv_tile(tile, nt, ga.nm, ga.bsame, m, ti, tj);
  const cx<T>* A = (const cx<T>*)g.A;
  const cx<T>* B = ((ga.bsame >> m) & 1) ? A : (const cx<T>*)g.B[m];
  const int64_t K = ga.K;
  const int64_t k0 = (int64_t)chunk * ga.kc, k1 = min(K, k0 + ga.kc);
  int ri[GV_T], rj[GV_T];
#pragma unroll
  for (int u = 0; u < GV_T; ++u) {
    ri[u] = min(ti * GV_T + u, nst - 1);  // clamped rows are computed and discarded
    rj[u] = min(tj * GV_T + u, nst - 1);
  }
  double cr[GV_T][GV_T], ci[GV_T][GV_T];
#pragma unroll
  for (int a = 0; a < GV_T; ++a)
#pragma unroll
    for (int b = 0; b < GV_T; ++b) cr[a][b] = ci[a][b] = 0.0;
  for (int64_t k = k0 + lane; k < k1; k += 64) {
    cx<double> av[GV_T], bv[GV_T];
#pragma unroll
    for (int u = 0; u < GV_T; ++u) {
      av[u] = ld64(A + (int64_t)ri[u] * K + k);
      bv[u] = ld64(B + (int64_t)rj[u] * K + k);
    }
#pragma unroll
    for (int a = 0; a < GV_T; ++a)
#pragma unroll
      for (int b = 0; b < GV_T; ++b) {  // conj(a) b
        cr[a][b] = fma(av[a].x, bv[b].x, fma(av[a].y, bv[b].y, cr[a][b]));
        ci[a][b] = fma(av[a].x, bv[b].y, fma(-av[a].y, bv[b].x, ci[a][b]));
      }
  }
#pragma unroll
  for (int a = 0; a < GV_T; ++a)
#pragma unroll
    for (int b = 0; b < GV_T; ++b)
      for (int o = 32; o > 0; o >>= 1) {
        cr[a][b] += __shfl_xor(cr[a][b], o, 64);
        ci[a][b] += __shfl_xor(ci[a][b], o, 64);
      }
  if (lane == 0) {
    const int np = nt * GV_T;
#pragma unroll
    for (int a = 0; a < GV_T; ++a)
#pragma unroll
      for (int b = 0; b < GV_T; ++b) {
        const int row = ti * GV_T + a, col = tj * GV_T + b;
        double* dst = g.part + 2 * ((((int64_t)m * np + row) * np + col) * ga.nchunk + chunk);
        dst[0] = cr[a][b];
        dst[1] = ci[a][b];
      }
  }
}

// Sum the chunks of each output element (one wavefront per element, fixed
// order); entries below the diagonal of a B == A matrix are the conjugates
// of the computed upper ones.
__global__ void __launch_bounds__(256) gram_vec_reduce_kernel(const double* __restrict__ part, int nchunk, int nm,
                                                              int np, int nst, int bsame, double scale,
                                                              double* __restrict__ out) {
  const int64_t total = (int64_t)nm * nst * nst;
  const int lane = threadIdx.x & 63;
  const int64_t wid = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t e = wid; e < total; e += nw) {
    const int m = e / ((int64_t)nst * nst);
    const int ij = e % ((int64_t)nst * nst);
    int i = ij / nst, j = ij % nst;
    const bool mirror = ((bsame >> m) & 1) && i > j;
    if (mirror) {
      const int t = i;
      i = j;
      j = t;
    }
    const double* p = part + 2 * (((int64_t)m * np + i) * np + j) * nchunk;
    double re = 0.0, im = 0.0;
    for (int ch = lane; ch < nchunk; ch += 64) {
      re += p[2 * ch];
      im += p[2 * ch + 1];
    }
    for (int o = 32; o > 0; o >>= 1) {
      re += __shfl_down(re, o, 64);
      im += __shfl_down(im, o, 64);
    }
    if (lane == 0) {
      out[2 * e] = re * scale;
      out[2 * e + 1] = (mirror ? -im : im) * scale;
    }
  }
}

// Rotation of a small block (nst <= 32) on the vector units: each lane owns
// one k, reads psi_i[k] for every i into registers, then writes
// out1_j[k] = sum_i psi_i[k] U1[i][j] (and out2 with U2).  A whole column is
// read before anything is written, so out1 may alias psi.  U is uniform
// across the wavefront and is read through the scalar cache.
template <typename T, int NI>
__global__ void __launch_bounds__(256) rotate_vec_kernel(const cx<T>* psi, const double* __restrict__ U1,
                                                         const double* __restrict__ U2, cx<T>* out1, cx<T>* out2,
                                                         int nst, int64_t K) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= K) return;
  cx<double> v[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) v[i] = i < nst ? ld64(psi + (int64_t)i * K + k) : cx<double>{0.0, 0.0};
  const cx<double>* u1 = (const cx<double>*)U1;
  const cx<double>* u2 = (const cx<double>*)U2;
  for (int j = 0; j < nst; ++j) {
    double r1 = 0.0, i1 = 0.0, r2 = 0.0, i2 = 0.0;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      if (i < nst) {
        const cx<double> a = u1[i * nst + j];
        r1 = fma(v[i].x, a.x, fma(-v[i].y, a.y, r1));
        i1 = fma(v[i].x, a.y, fma(v[i].y, a.x, i1));
        if (u2) {
          const cx<double> b = u2[i * nst + j];
          r2 = fma(v[i].x, b.x, fma(-v[i].y, b.y, r2));
          i2 = fma(v[i].x, b.y, fma(v[i].y, b.x, i2));
        }
      }
    }
    out1[(int64_t)j * K + k] = cx<T>{T(r1), T(i1)};
    if (u2) out2[(int64_t)j * K + k] = cx<T>{T(r2), T(i2)};
  }
}

}  // namespace

extern "C" int hfb_gram_workspace(int nst, long long K) {
  const int np = (nst + 15) & ~15;
  const int mfma = gram_nchunk(np, K) * 4 * np * np * 2;  // doubles for up to four matrices
  const int np4 = (nst + GV_T - 1) / GV_T * GV_T;
  const int vec = gram_vec_nchunk(K) * 4 * np4 * np4 * 2;
  return mfma > vec ? mfma : vec;
}

template <typename T, int NM>
int launch_gram(const GramArgs& g, int npmax, hipStream_t st) {
  const int ntile = (npmax / 16) * (npmax / 16) * NM;
  const size_t lds = sizeof(cx<double>) * (size_t)npmax * LDR * (1 + NM);
  const dim3 grid(g.nchunk, 2);
  const int ne = (npmax * KS + GRAM_NT - 1) / GRAM_NT;
#define HFB_GR(MT, NEV)                                                          \
  {                                                                              \
    auto kern = gram_kernel<T, NM, MT, NEV>;                                     \
    if (lds > 32 * 1024) allow_full_lds((const void*)kern);                      \
    hipLaunchKernelGGL(kern, grid, dim3(GRAM_NT), lds, st, g);                  \
  }
  const int per_wave = (ntile + 3) / 4;
  if (ne > 8) return HFB_E_SIZE;
#define HFB_GRN(MT)                    \
  if (ne <= 2) HFB_GR(MT, 2)           \
  else if (ne <= 4) HFB_GR(MT, 4)      \
  else HFB_GR(MT, 8)
  if (per_wave <= 1) { HFB_GRN(1) }
  else if (per_wave <= 2) { HFB_GRN(2) }
  else if (per_wave <= 4) { HFB_GRN(4) }
  else if (per_wave <= 8) { HFB_GRN(8) }
  else if (per_wave <= 16) { HFB_GRN(16) }
  else return HFB_E_SIZE;
#undef HFB_GRN
#undef HFB_GR
  return launch_status();
}

// Both isospin blocks in one launch: A[iq], B[iq*4 + m], out[iq] (nm, nst_iq,
// nst_iq) complex, part[iq] workspaces of hfb_gram_workspace(nst_iq, K).
// A block with nst 0 is skipped.
extern "C" int hfb_gram2(int prec, const void* const* A, const void* const* B, int nm, const int* nst, long long K,
                         double scale, double* const* part, double* const* out, void* stream) {
  if (!A || !B || !part || !out || !nst || nm < 1 || nm > 4 || K <= 0) return HFB_E_ARG;
  GramArgs g{};
  g.bsame = 0;
  for (int m = 0; m < nm; ++m) {
    bool same = true;
    for (int iq = 0; iq < 2; ++iq)
      if (nst[iq] > 0 && B[iq * 4 + m] != A[iq]) same = false;
    if (same) g.bsame |= 1 << m;
  }
  int npmax = 0;
  for (int iq = 0; iq < 2; ++iq) {
    GramIso& gi = g.iso[iq];
    gi.nst = nst[iq];
    gi.np = nst[iq] > 0 ? (nst[iq] + 15) & ~15 : 0;
    if (nst[iq] > 0) {
      if (!A[iq] || !part[iq] || !out[iq]) return HFB_E_ARG;
      gi.A = A[iq];
      for (int m = 0; m < nm; ++m) {
        if (!B[iq * 4 + m]) return HFB_E_ARG;
        gi.B[m] = B[iq * 4 + m];
      }
      gi.part = part[iq];
    }
    npmax = gi.np > npmax ? gi.np : npmax;
  }
  if (npmax == 0) return HFB_OK;
  if (npmax <= 64) {  // small blocks: vector FP64 path
    GramVecArgs v{};
    v.iso[0] = g.iso[0];
    v.iso[1] = g.iso[1];
    v.nm = nm;
    v.bsame = g.bsame;
    v.K = K;
    v.nchunk = gram_vec_nchunk(K);
    v.kc = (K + v.nchunk - 1) / v.nchunk;
    int items = 0;
    for (int iq = 0; iq < 2; ++iq) {
      const int nt = (nst[iq] + GV_T - 1) / GV_T;
      int cnt = 0;
      for (int m = 0; m < nm; ++m) cnt += ((v.bsame >> m) & 1) ? nt * (nt + 1) / 2 : nt * nt;
      v.ntile[iq] = nst[iq] > 0 ? cnt : 0;
      items = v.ntile[iq] * v.nchunk > items ? v.ntile[iq] * v.nchunk : items;
    }
    hipStream_t st = (hipStream_t)stream;
    if (prec == HFB_F64)
      hipLaunchKernelGGL(gram_vec_kernel<double>, dim3((items + 3) / 4, 2), dim3(256), 0, st, v);
    else if (prec == HFB_F32)
      hipLaunchKernelGGL(gram_vec_kernel<float>, dim3((items + 3) / 4, 2), dim3(256), 0, st, v);
    else
      return HFB_E_ARG;
    int rc = launch_status();
    for (int iq = 0; iq < 2 && rc == HFB_OK; ++iq) {
      if (nst[iq] <= 0) continue;
      const int np4 = (nst[iq] + GV_T - 1) / GV_T * GV_T;
      const int64_t tot = (int64_t)nm * nst[iq] * nst[iq];
      int nb = (int)((tot * 64 + 255) / 256);
      if (nb > 4096) nb = 4096;
      hipLaunchKernelGGL(gram_vec_reduce_kernel, dim3(nb < 1 ? 1 : nb), dim3(256), 0, st, part[iq], v.nchunk, nm,
                         np4, nst[iq], v.bsame, scale, out[iq]);
      rc = launch_status();
    }
    return rc;
  }
  g.nm = nm;
  g.K = K;
  g.nchunk = gram_nchunk(npmax, K);
  int64_t kc = (K + g.nchunk - 1) / g.nchunk;
  kc = (kc + KS - 1) / KS * KS;
  g.kc = kc;
  hipStream_t st = (hipStream_t)stream;
  int rc;
#define HFB_GT(TT)                                          \
  if (nm == 1) rc = launch_gram<TT, 1>(g, npmax, st);       \
  else if (nm == 2) rc = launch_gram<TT, 2>(g, npmax, st);  \
  else if (nm == 3) rc = launch_gram<TT, 3>(g, npmax, st);  \
  else rc = launch_gram<TT, 4>(g, npmax, st);
  if (prec == HFB_F64) {
    HFB_GT(double)
  } else if (prec == HFB_F32) {
    HFB_GT(float)
  } else {
    return HFB_E_ARG;
  }
#undef HFB_GT
  if (rc != HFB_OK) return rc;
  for (int iq = 0; iq < 2 && rc == HFB_OK; ++iq) {
    if (nst[iq] <= 0) continue;
    const int64_t tot = (int64_t)nm * nst[iq] * nst[iq];
    int nb = (int)((tot * 64 + 255) / 256);
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(gram_reduce_kernel, dim3(nb < 1 ? 1 : nb), dim3(256), 0, st, part[iq], g.nchunk, nm,
                       g.iso[iq].np, nst[iq], scale, out[iq]);
    rc = launch_status();
  }
  return rc;
}

// One isospin block (kept for callers of a single block).
extern "C" int hfb_gram(int prec, const void* A, const void* const* B, int nm, int nst, long long K, double scale,
                        double* part, double* out, void* stream) {
  if (!B) return HFB_E_ARG;
  const void* As[2] = {A, nullptr};
  const void* Bs[8] = {nullptr};
  for (int m = 0; m < nm && m < 4; ++m) Bs[m] = B[m];
  const int ns[2] = {nst, 0};
  double* parts[2] = {part, nullptr};
  double* outs[2] = {out, nullptr};
  return hfb_gram2(prec, As, Bs, nm, ns, K, scale, parts, outs, stream);
}

extern "C" int hfb_rotate(int prec, const void* psi, const double* U1, const double* U2, void* out1, void* out2,
                          int nst, long long K, void* stream) {
  if (!psi || !U1 || !out1 || nst <= 0 || K <= 0) return HFB_E_ARG;
  if (U2 && !out2) return HFB_E_ARG;
  if (nst <= 32) {  // small blocks: register columns on the vector units
    const int nb = (int)((K + 255) / 256);
    hipStream_t st0 = (hipStream_t)stream;
#define HFB_RV(TT, NIV)                                                                                   \
  hipLaunchKernelGGL((rotate_vec_kernel<TT, NIV>), dim3(nb), dim3(256), 0, st0, (const cx<TT>*)psi, U1, U2, \
                     (cx<TT>*)out1, (cx<TT>*)out2, nst, (int64_t)K)
    if (prec == HFB_F64) {
      if (nst <= 8) HFB_RV(double, 8);
      else if (nst <= 16) HFB_RV(double, 16);
      else if (nst <= 24) HFB_RV(double, 24);
      else HFB_RV(double, 32);
    } else if (prec == HFB_F32) {
      if (nst <= 8) HFB_RV(float, 8);
      else if (nst <= 16) HFB_RV(float, 16);
      else if (nst <= 24) HFB_RV(float, 24);
      else HFB_RV(float, 32);
    } else {
      return HFB_E_ARG;
    }
#undef HFB_RV
    return launch_status();
  }
  const int np = (nst + 15) & ~15;
  const int ntj = np / 16;
  const int njt = ntj < 4 ? ntj : 4;  // j-tiles per workgroup column
  const int jgroups = (ntj + njt - 1) / njt;
  if (jgroups > 1 && out1 == psi) return HFB_E_ARG;  // in place only when one column of blocks covers all j
  if (ntj % njt != 0) return HFB_E_SIZE;
  const size_t lds = sizeof(double) * 2 * 2 * (size_t)np * np;
  if (lds > 152 * 1024) return HFB_E_SIZE;
  const int64_t nblk = (K + 15) / 16;
  int gx = (int)((nblk + 4 * 2 - 1) / (4 * 2));  // two 16-point blocks per wavefront
  if (gx < 1) gx = 1;
  hipStream_t st = (hipStream_t)stream;
#define HFB_RT(TT, NJ)                                                                                          \
  {                                                                                                             \
    auto kern = rotate_kernel<TT, NJ>;                                                                          \
    if (lds > 32 * 1024) allow_full_lds((const void*)kern);                                                     \
    hipLaunchKernelGGL(kern, dim3(gx, jgroups), dim3(256), lds, st, (const cx<TT>*)psi, U1, U2, (cx<TT>*)out1,    \
                       (cx<TT>*)out2, nst, (int64_t)K);                                                         \
  }
#define HFB_RTJ(TT)                 \
  if (njt == 1) HFB_RT(TT, 1)       \
  else if (njt == 2) HFB_RT(TT, 2)  \
  else if (njt == 3) HFB_RT(TT, 3)  \
  else HFB_RT(TT, 4)
  if (prec == HFB_F64) {
    HFB_RTJ(double)
  } else if (prec == HFB_F32) {
    HFB_RTJ(float)
  } else {
    return HFB_E_ARG;
  }
#undef HFB_RTJ
#undef HFB_RT
  return launch_status();
}
