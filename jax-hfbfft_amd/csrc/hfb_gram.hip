// Subspace matrices and rotations of diagstep (static.py:261-577).
//
// Gram: C_m = A^H B_m for up to four right-hand blocks sharing A = psi
// (overlap psi^H psi, H_old = psi^H hmfpsi, lambda_lin = psi^H hampsi,
// G_old = psi^H delpsi), K = 2*nx*ny*nz per state.
//
// Rotation: out_j = sum_i psi_i U[i][j] for two matrices at once (psi U and
// psi (U lambda+), i.e. the rotated states and the Lagrange field).
//
// FP32 blocks (mixed precision) run on the FP32 matrix cores: the Gram on
// 4x4 tiles (no padding beyond a multiple of 4 states), the rotation on
// 16x16 tiles (up to 24 states with the whole coefficient matrix in
// registers, above that one 16-row tile per wavefront and k-step).  FP64
// blocks run as register-tiled FMAs on the vector units.
#include <math.h>

#include "hfb_ctx.h"
#include "hfb_jobs.h"

using namespace hfb;


namespace {


template <typename T>
__device__ __forceinline__ cx<double> ld64(const cx<T>* p) {
  const cx<T> v = *p;
  return {double(v.x), double(v.y)};
}

struct GramIso {
  const void* A;
  const void* B[4];
  int nst;
  int np;       // nst padded to the tile edge
  double* part; // [m][np][np][chunk] complex
};


// ---------------------------------------------------------------------------
// Gram on the vector FP64 units: a wavefront owns a 4x4 tile of one C_m
// (complex, in registers) and a chunk of K; lanes stride over k (coalesced
// rows); the lane
// sums are combined by shuffles and the chunks by gram_vec_reduce_kernel in
// a fixed order.  For B_m == A only tiles with ti <= tj are computed; the
// reduction mirrors the rest (C^H = C).
// ---------------------------------------------------------------------------
constexpr int GV_T = 4;      // tile edge
constexpr int GV_KPL = 32;   // k per lane per chunk
constexpr int GV_CHUNK = 64 * GV_KPL;

// K-chunks: at least GV_KPL elements of K per lane, and about 16K
// wavefront items over the tiles of one matrix (fewer chunks for big blocks)
int gram_vec_nchunk(int nst, int64_t K) {
  const int nt = (nst + GV_T - 1) / GV_T;
  int64_t nc = (K + GV_CHUNK - 1) / GV_CHUNK;
  int64_t cap = 16384 / ((int64_t)nt * nt);
  if (cap < 4) cap = 4;
  if (nc > cap) nc = cap;
  return nc < 1 ? 1 : (int)nc;
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
__global__ void __launch_bounds__(256, 3) gram_vec_kernel(GramVecArgs ga, hfb_finish_job fjob) {
  if (fjob.nstates > 0 && blockIdx.x == gridDim.x - 1) {  // the extra column: the step's scalar tail
    if (blockIdx.y == 0) grstep_finish_body(fjob);
    return;
  }
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
#pragma unroll 2
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
  // lane sums, level by level over all 32 values so the chains overlap
#define HFB_GV_LEVEL(XOR)                                \
  _Pragma("unroll") for (int a = 0; a < GV_T; ++a)       \
  _Pragma("unroll") for (int b = 0; b < GV_T; ++b) {     \
    cr[a][b] += XOR(cr[a][b]);                           \
    ci[a][b] += XOR(ci[a][b]);                           \
  }
  HFB_GV_LEVEL(lane_xor<1>)
  HFB_GV_LEVEL(lane_xor<2>)
  HFB_GV_LEVEL(lane_xor<4>)
  HFB_GV_LEVEL(lane_xor<8>)
  HFB_GV_LEVEL(lane_xor<16>)
#undef HFB_GV_LEVEL
#pragma unroll
  for (int a = 0; a < GV_T; ++a)
#pragma unroll
    for (int b = 0; b < GV_T; ++b) {
      cr[a][b] += __shfl_xor(cr[a][b], 32, 64);
      ci[a][b] += __shfl_xor(ci[a][b], 32, 64);
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
struct GramRedIso {
  const double* part;
  double* out;
  int nst, np;
};

// Chunk sums of both isospins (blockIdx.y), one wavefront per entry.
__global__ void __launch_bounds__(256) gram_vec_reduce_kernel(GramRedIso r0, GramRedIso r1, int nchunk, int nm,
                                                              int bsame, double scale) {
  const GramRedIso& r = blockIdx.y == 0 ? r0 : r1;
  const double* __restrict__ part = r.part;
  double* __restrict__ out = r.out;
  const int nst = r.nst, np = r.np;
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

// ---------------------------------------------------------------------------
// Gram of FP32 blocks on the matrix cores (v_mfma_f32_4x4x1_16b_f32: 16
// independent 4x4x1 outer products, "blocks", per instruction).  In real form
// conj(a) b = (ar br + ai bi) + i (ar bi - ai br) over the interleaved
// (re, im) floats of a state row.  A wavefront owns up to GM_BT x GM_BT
// tiles of 4 x 4 entries of one C_m (rows of A: tiles ti0.., of B: tj0..) and
// a chunk of the row; per group of 64 floats (32 complex points) block b
// takes floats 4b..4b+3: lane l = 4 b + r loads that float4 of row 4 t + r
// (16 lanes read one 256-byte row segment) and its four components enter
// four products of the block, (re, im, re, im) against the same of B for the
// real part and against (im, -re, im, -re) for the imaginary part.  Every
// block accumulates its slices of the chunk in FP32 (at most GM_GMAX groups,
// 4 GM_GMAX products); at the end the 16 block sums of each entry are added
// through LDS (one row of tiles per pass) and written as FP64 chunk partials
// in gram_vec_kernel's layout, summed by gram_vec_reduce_kernel.
// Operand layout of the instruction: A lane 4b+i = A_b[i], B lane 4b+j =
// B_b[j], result register i of lane 4b+j = D_b[i][j].
// ---------------------------------------------------------------------------
typedef float gm_v4 __attribute__((ext_vector_type(4)));
constexpr int GM_BT = 5;     // tiles per block edge at most (2 x 25 accumulators fit the AGPRs)
constexpr int GM_GMAX = 64;  // 64-float groups per chunk at most (FP32 sums of 256 products)
constexpr int GM_GMIN = 8;   // and at least, where the row allows

struct GramMfmaIso {
  int nst, nt, nb;  // states, 4-row tiles, blocks per edge
  int nsub;         // (matrix, block) items per chunk
};

struct GramMfmaArgs {
  GramIso iso[2];
  GramMfmaIso mi[2];
  int nm, bsame;
  int64_t Kr;    // floats per state row
  int gpc;       // groups per chunk
  int ngroup;    // groups per row
  int nchunk;
  int64_t nitem; // nchunk * (nsub0 + nsub1)
};

// Tiles [lo, hi) of block i of nb over nt tiles (balanced).
__device__ __forceinline__ void gm_span(int i, int nb, int nt, int& lo, int& cnt) {
  lo = i * nt / nb;
  cnt = (i + 1) * nt / nb - lo;
}

// (matrix, block row, block column) of sub-item s: matrices in order; for a
// matrix with B == A only blocks bi <= bj.
__device__ __forceinline__ void gm_sub(int s, int nb, int nm, int bsame, int& m, int& bi, int& bj) {
  m = 0;
  for (; m < nm; ++m) {
    const int cnt = ((bsame >> m) & 1) ? nb * (nb + 1) / 2 : nb * nb;
    if (s < cnt) break;
    s -= cnt;
  }
  if (!((bsame >> m) & 1)) {
    bi = s / nb;
    bj = s % nb;
    return;
  }
  bi = 0;
  while (s >= nb - bi) {
    s -= nb - bi;
    ++bi;
  }
  bj = bi + s;
}

// The eight products of every tile for one group, component by component
// over all tiles, so that an accumulator's next product comes 2 BT^2
// instructions later (no waiting on the matrix pipeline).  Tiles of a
// partial block read clamped rows: computed, never summed.
template <int BT>
__device__ __forceinline__ void gm_mma(gm_v4 (&accr)[BT][BT], gm_v4 (&acci)[BT][BT], const gm_v4 (&a)[BT],
                                       const gm_v4 (&b)[BT]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
#pragma unroll
    for (int u = 0; u < BT; ++u) {
      // real part: (re, im) x (re, im); imaginary part: x (im, -re)
      const float br = b[u][c];
      const float bi = (c & 1) ? -b[u][c - 1] : b[u][c + 1];
#pragma unroll
      for (int t = 0; t < BT; ++t) {
        accr[t][u] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[t][c], br, accr[t][u], 0, 0, 0);
        acci[t][u] = __builtin_amdgcn_mfma_f32_4x4x1f32(a[t][c], bi, acci[t][u], 0, 0, 0);
      }
    }
  }
}

// BT: tiles per block edge of this launch (at most GM_BT).  Groups go from
// HBM straight into a ring of GM_RING LDS slots per wavefront
// (global_load_lds, load i of a group fills 64 float4 in lane order, so a
// lane reads back its own float4), GM_RING - 1 groups ahead of their
// products: the one wavefront per SIMD has two groups of products to cover
// a load, and the waits are explicit (vm_wait) rather than the compiler's
// count across the loop.
constexpr int GM_RING = 3;

template <int BT>
constexpr size_t gram_mfma_lds() { return (size_t)4 * GM_RING * 2 * BT * 64 * sizeof(gm_v4); }

template <int BT>
__global__ void __launch_bounds__(256, 1) gram_mfma_kernel(GramMfmaArgs ga, hfb_finish_job fjob) {
  if (fjob.nstates > 0 && blockIdx.x == gridDim.x - 1) {  // the extra workgroup: the step's scalar tail
    grstep_finish_body(fjob);
    return;
  }
  extern __shared__ __attribute__((aligned(16))) unsigned char gm_lds[];
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // this wavefront's ring; after the loop its first 2 BT x 64 float4 hold
  // one row of tiles of the block sums
  gm_v4* const ring = reinterpret_cast<gm_v4*>(gm_lds) + (size_t)wv * GM_RING * 2 * BT * 64;
  const int64_t item = (int64_t)blockIdx.x * 4 + wv;
  const bool live = item < ga.nitem;
  int iq = 0, m = 0, ti0 = 0, nti = 0, tj0 = 0, ntj = 0, chunk = 0;
  if (live) {
    const int nsub = ga.mi[0].nsub + ga.mi[1].nsub;
    chunk = (int)(item / nsub);
    int s = (int)(item % nsub);
    if (s >= ga.mi[0].nsub) {
      s -= ga.mi[0].nsub;
      iq = 1;
    }
    int bi, bj;
    gm_sub(s, ga.mi[iq].nb, ga.nm, ga.bsame, m, bi, bj);
    gm_span(bi, ga.mi[iq].nb, ga.mi[iq].nt, ti0, nti);
    gm_span(bj, ga.mi[iq].nb, ga.mi[iq].nt, tj0, ntj);
  }
  iq = __builtin_amdgcn_readfirstlane(iq);
  m = __builtin_amdgcn_readfirstlane(m);
  ti0 = __builtin_amdgcn_readfirstlane(ti0);
  nti = __builtin_amdgcn_readfirstlane(nti);
  tj0 = __builtin_amdgcn_readfirstlane(tj0);
  ntj = __builtin_amdgcn_readfirstlane(ntj);
  chunk = __builtin_amdgcn_readfirstlane(chunk);
  const GramIso& g = ga.iso[iq];
  const int nst = ga.mi[iq].nst;
  const bool sym = (ga.bsame >> m) & 1;
  const char* A = (const char*)g.A;
  const char* B = sym ? A : (const char*)g.B[m];
  const int r = lane & 3, blk = lane >> 2;
  // byte offsets of this lane's rows (rows past nst clamped: computed, never
  // summed); a block is below 4 GiB (checked by the host)
  uint32_t aoff[BT], boff[BT];
#pragma unroll
  for (int t = 0; t < BT; ++t) {
    aoff[t] = (uint32_t)(((int64_t)min(4 * (ti0 + t) + r, nst - 1) * ga.Kr + 4 * blk) * 4);
    boff[t] = (uint32_t)(((int64_t)min(4 * (tj0 + t) + r, nst - 1) * ga.Kr + 4 * blk) * 4);
  }
  gm_v4 accr[BT][BT], acci[BT][BT];
#pragma unroll
  for (int t = 0; t < BT; ++t)
#pragma unroll
    for (int u = 0; u < BT; ++u) accr[t][u] = acci[t][u] = gm_v4{0.f, 0.f, 0.f, 0.f};
  if (live) {
    const int g0 = chunk * ga.gpc, g1 = min(ga.ngroup, g0 + ga.gpc);
    // group gg into ring slot k: 2 BT loads, unconditional (groups past the
    // chunk's end load the last group again, never used), so every group
    // is the same number of loads in flight
    auto stage = [&](int k, int gg) {
      const int64_t go = (int64_t)min(gg, g1 - 1) * 256;
      gm_v4* dst = ring + k * 2 * BT * 64;
#pragma unroll
      for (int t = 0; t < BT; ++t)
        __builtin_amdgcn_global_load_lds((const void*)(A + go + aoff[t]), (lds_void)(dst + t * 64), 16, 0, 0);
#pragma unroll
      for (int u = 0; u < BT; ++u)
        __builtin_amdgcn_global_load_lds((const void*)(B + go + boff[u]), (lds_void)(dst + (BT + u) * 64), 16, 0,
                                         0);
    };
    static_assert(GM_RING == 3, "the loop below runs the ring's three slots");
    stage(0, g0);
    stage(1, g0 + 1);
    stage(2, g0 + 2);
    auto step = [&](int k, int gi) {
      vm_wait<(GM_RING - 1) * 2 * BT>();  // slot k has landed; the two younger groups stay in flight
      const gm_v4* src = ring + k * 2 * BT * 64;
      gm_v4 a[BT], b[BT];
#pragma unroll
      for (int t = 0; t < BT; ++t) a[t] = src[t * 64 + lane];
#pragma unroll
      for (int u = 0; u < BT; ++u) b[u] = src[(BT + u) * 64 + lane];
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): the slot is read before it is loaded again
      asm volatile("" ::: "memory");
      stage(k, gi + GM_RING);
      if (gi < g1) gm_mma<BT>(accr, acci, a, b);
    };
#pragma unroll 1
    for (int gi = g0; gi < g1; gi += 3) {
      step(0, gi);
      step(1, gi + 1);
      step(2, gi + 2);
    }
    vm_wait<0>();  // nothing in flight into the ring before it holds the block sums
  }
  // block sums through this wavefront's ring (idle now; only this
  // wavefront touches it, and its LDS accesses run in order: no barrier),
  // GM_TP rows of tiles per pass: lane 4b+j stores registers 0..3 of tile
  // (t, u) of each part at R[b][(t - t0, part, u, j)] (a float4 per tile);
  // lane (t - t0, u, j) then adds the 16 float4 of both parts (consecutive
  // lanes read consecutive float4) and writes entries (4 t + i, 4 u + j).
  constexpr int GM_TP = 3;  // rows of tiles per pass: 16 blocks x 3 x 2 parts x BT x 4 fill the ring
  static_assert(16 * GM_TP * 2 * BT * 4 <= GM_RING * 2 * BT * 64, "block sums fit the ring");
  static_assert(GM_TP * BT * 4 <= 64, "one lane per (row, tile, column) of a pass");
  constexpr int SB = GM_TP * 2 * BT * 4;  // float4 per block
  gm_v4* R = ring;
  const int j = lane & 3;
  const int np = g.np;
#pragma unroll
  for (int t0 = 0; t0 < BT; t0 += GM_TP) {
#pragma unroll
    for (int tt = 0; tt < GM_TP; ++tt) {
      if (t0 + tt >= BT) break;
#pragma unroll
      for (int u = 0; u < BT; ++u) {
        R[blk * SB + ((tt * 2 + 0) * BT + u) * 4 + j] = accr[t0 + tt][u];
        R[blk * SB + ((tt * 2 + 1) * BT + u) * 4 + j] = acci[t0 + tt][u];
      }
    }
    const int tt = lane / (4 * BT), u = (lane >> 2) % BT;
    const int t = t0 + tt;
    if (live && lane < GM_TP * BT * 4 && t < BT && t < nti && u < ntj) {
      gm_v4 sr = R[((tt * 2 + 0) * BT + u) * 4 + j];
      gm_v4 si = R[((tt * 2 + 1) * BT + u) * 4 + j];
#pragma unroll
      for (int bb = 1; bb < 16; ++bb) {
        sr += R[bb * SB + ((tt * 2 + 0) * BT + u) * 4 + j];
        si += R[bb * SB + ((tt * 2 + 1) * BT + u) * 4 + j];
      }
      const int col = 4 * (tj0 + u) + j;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 4 * (ti0 + t) + i;
        double* dst = g.part + 2 * ((((int64_t)m * np + row) * np + col) * ga.nchunk + chunk);
        dst[0] = (double)sr[i];
        dst[1] = (double)si[i];
      }
    }
  }
}

// Compute units of the current device (cached per device).
int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] <= 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    cus[dev] = v;
  }
  return cus[dev];
}

// Chunks of the matrix-core Gram: at most GM_GMAX groups each, at least
// GM_GMIN where the row allows, otherwise as many as fit one round of the
// device: the kernel runs one workgroup of four wavefronts per compute unit,
// so at most 4 (CUs - 1) wavefronts over nsub (matrix, block) items, one
// compute unit left for the extra workgroup of the step's tail (measured at
// Ca40: 33 us for the 950 wavefronts of one round against 45 us for 1536,
// and 44 us when two workgroups spill into a second round).
int gram_mfma_nchunk(int64_t ngroup, int nsub) {
  const int64_t waves = 4 * (int64_t)(device_cus() - 1);
  int64_t nc = waves / (nsub > 0 ? nsub : 1);
  if (nc < 1) nc = 1;
  const int64_t most = ngroup / GM_GMIN > 0 ? ngroup / GM_GMIN : 1;
  if (nc > most) nc = most;
  const int64_t least = (ngroup + GM_GMAX - 1) / GM_GMAX;
  if (nc < least) nc = least;
  return (int)nc;
}

int gram_mfma_blocks(int nst) {
  const int nt = (nst + 3) / 4;
  return (nt + GM_BT - 1) / GM_BT;
}

// Rotation out_j[k] = sum_i psi_i[k] U[i][j] (two matrices U1, U2 at
// once).  A workgroup stages psi_i[k] for every i and its 64*KPL
// consecutive k in LDS; each wavefront then produces outputs j in groups of
// RJ for those k, KPL values of k per lane.  The group index is made
// wavefront-uniform, so U1[i][j], U2[i][j] are read through the scalar
// cache into scalar registers (no LDS bandwidth spent on broadcasts).  A
// workgroup reads all of its columns of psi before its first store, so out1
// may alias psi.
struct RotIso {
  const void* psi;
  const double* U1;
  const double* U2;  // may be null (then out2 is not written)
  void* out1;
  void* out2;
  int nst;
};

struct RotArgs {
  RotIso iso[2];
  int64_t K;
};

template <typename T, int RJ, int KPL>
__global__ void __launch_bounds__(256) rotate_kernel(RotArgs ra) {
  const RotIso& ri = ra.iso[blockIdx.y];
  const int nst = ri.nst;
  const int64_t K = ra.K;
  const cx<T>* psi = (const cx<T>*)ri.psi;
  const double* __restrict__ U1 = ri.U1;
  const double* __restrict__ U2 = ri.U2;
  cx<T>* out1 = (cx<T>*)ri.out1;
  cx<T>* out2 = (cx<T>*)ri.out2;
  constexpr int KT = 64 * KPL;  // k per workgroup
  extern __shared__ unsigned char rot_lds[];
  cx<T>* sp = (cx<T>*)rot_lds;  // [nst][KT]
  const bool two = U2 != nullptr;
  const int64_t k0 = (int64_t)blockIdx.x * KT;
  if (nst <= 0 || k0 >= K) return;  // uniform over the workgroup
  for (int e = threadIdx.x; e < nst * KT; e += blockDim.x) {
    const int i = e / KT, kk = e % KT;
    const int64_t k = k0 + kk;
    sp[e] = k < K ? psi[(int64_t)i * K + k] : cx<T>{T(0), T(0)};
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int jg = wave * RJ; jg < nst; jg += 4 * RJ) {
    // the last group is moved down to end at nst (its first outputs are then
    // also produced, with identical values, by the previous group), so the
    // RJ coefficients of a row are contiguous and load as one scalar block
    const int j0 = min(jg, nst - RJ);
    double r1[KPL][RJ], i1[KPL][RJ], r2[KPL][RJ], i2[KPL][RJ];
#pragma unroll
    for (int c = 0; c < KPL; ++c)
#pragma unroll
      for (int t = 0; t < RJ; ++t) r1[c][t] = i1[c][t] = r2[c][t] = i2[c][t] = 0.0;
    for (int i = 0; i < nst; ++i) {
      double vx[KPL], vy[KPL];
#pragma unroll
      for (int c = 0; c < KPL; ++c) {
        const cx<T> pv = sp[i * KT + lane + 64 * c];
        vx[c] = double(pv.x);
        vy[c] = double(pv.y);
      }
      const double* u1 = U1 + 2 * ((int64_t)i * nst + j0);
      const double* u2 = two ? U2 + 2 * ((int64_t)i * nst + j0) : u1;
      double ax[RJ], ay[RJ], bx[RJ], by[RJ];
#pragma unroll
      for (int t = 0; t < RJ; ++t) {
        ax[t] = u1[2 * t];
        ay[t] = u1[2 * t + 1];
        bx[t] = u2[2 * t];
        by[t] = u2[2 * t + 1];
      }
#pragma unroll
      for (int t = 0; t < RJ; ++t) {
#pragma unroll
        for (int c = 0; c < KPL; ++c) {
          r1[c][t] = fma(vx[c], ax[t], fma(-vy[c], ay[t], r1[c][t]));
          i1[c][t] = fma(vx[c], ay[t], fma(vy[c], ax[t], i1[c][t]));
          r2[c][t] = fma(vx[c], bx[t], fma(-vy[c], by[t], r2[c][t]));
          i2[c][t] = fma(vx[c], by[t], fma(vy[c], bx[t], i2[c][t]));
        }
      }
    }
#pragma unroll
    for (int c = 0; c < KPL; ++c) {
      const int64_t k = k0 + lane + 64 * c;
      if (k >= K) continue;
#pragma unroll
      for (int t = 0; t < RJ; ++t) {
        const int j = j0 + t;
        out1[(int64_t)j * K + k] = cx<T>{T(r1[c][t]), T(i1[c][t])};
        if (two) out2[(int64_t)j * K + k] = cx<T>{T(r2[c][t]), T(i2[c][t])};
      }
    }
  }
}

// FP32 rotation on the matrix cores (v_mfma_f32_16x16x4f32) for blocks of
// at most 24 states.  As a real product, out = C psi with the 2 nst input
// columns (Re psi_i, Im psi_i) and the out columns (Re, Im of out1_j, then
// of out2_j) of both matrices stacked: C[(m, j, co)][(i, ci)] = U_m[i][j].x
// for ci == co, -U_m[i][j].y for (ci, co) = (1, 0), U_m[i][j].y for (0, 1).
// The workgroup copies U1, U2 (FP32) to LDS; each wavefront keeps C as
// MT x KS A fragments (rows: out columns, 16 per tile; k-steps: four input
// columns) and streams tiles of 16 grid points, TPW tiles loaded at a time:
// B fragment lane l = input column 4 ks + l / 16 at point k0 + l % 16, a
// 4-byte load straight from psi.  Result lane l, register r = out column
// 16 mt + 4 (l / 16) + r at point k0 + l % 16, so registers (0, 1) and
// (2, 3) are the (Re, Im) of one state and store as one 8-byte value, 16
// consecutive points per state and lane group.  A wavefront reads every
// input of its points before their first store, so out1 may alias psi.
typedef float rv4f __attribute__((ext_vector_type(4)));
constexpr int ROT_MFMA_NMAX = 24;
constexpr int ROT_TPW = 8;  // tiles of 16 points a wavefront loads at once
constexpr int ROT_BPC = 1;  // workgroups per CU

template <int KS, int MT, int TPW>
__global__ void __launch_bounds__(256) rotate_mfma_kernel(RotArgs ra) {
  __shared__ float2 Us[2][ROT_MFMA_NMAX * ROT_MFMA_NMAX];
  const RotIso& ri = ra.iso[blockIdx.y];
  const int nst = ri.nst;
  if (nst <= 0) return;  // uniform over the workgroup
  const int64_t K = ra.K;
  const bool two = ri.U2 != nullptr;
  const int ncol = (two ? 4 : 2) * nst;  // out columns of both matrices
  for (int e = threadIdx.x; e < nst * nst; e += blockDim.x) {
    Us[0][e] = float2{float(ri.U1[2 * e]), float(ri.U1[2 * e + 1])};
    if (two) Us[1][e] = float2{float(ri.U2[2 * e]), float(ri.U2[2 * e + 1])};
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, g = lane >> 4, c16 = lane & 15;
  float A[MT][KS];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int co_all = 16 * mt + c16;
    const bool ok = co_all < ncol;
    const int m = ok ? co_all / (2 * nst) : 0, j = ok ? (co_all % (2 * nst)) >> 1 : 0, co = co_all & 1;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int i = 2 * ks + (g >> 1), ci = g & 1;
      float v = 0.0f;
      if (ok && i < nst) {
        const float2 u = Us[m][i * nst + j];
        v = ci == co ? u.x : (ci == 1 ? -u.y : u.y);
      }
      A[mt][ks] = v;
    }
  }
  // this lane's out columns: state offset j K into out1 or out2 (or none)
  int64_t ooff[MT][2];
  int osel[MT][2];  // 0: out1, 1: out2, -1: none
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int co_all = 16 * mt + 4 * g + 2 * h;
      osel[mt][h] = co_all < ncol ? co_all / (2 * nst) : -1;
      ooff[mt][h] = (int64_t)((co_all % (2 * nst)) >> 1) * K;
    }
  const float* psi = (const float*)ri.psi;
  float2* o1 = (float2*)ri.out1;
  float2* o2 = (float2*)ri.out2;
  const int64_t ntile = K / 16;
  const int64_t nwave = (int64_t)gridDim.x * 4;
  for (int64_t base = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * TPW; base < ntile; base += nwave * TPW) {
    // the inputs of all TPW tiles are loaded before the first product
    float B[TPW][KS];
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int64_t k = (base + t) * 16 + c16;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int i = 2 * ks + (g >> 1);
        B[t][ks] = (base + t < ntile && i < nst) ? psi[2 * ((int64_t)i * K + k) + (g & 1)] : 0.0f;
      }
    }
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      if (base + t >= ntile) break;
      const int64_t k = (base + t) * 16 + c16;
      rv4f acc[MT];
#pragma unroll
      for (int mt = 0; mt < MT; ++mt) acc[mt] = rv4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
          acc[mt] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mt][ks], B[t][ks], acc[mt], 0, 0, 0);
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (osel[mt][h] < 0) continue;
          float2* o = osel[mt][h] == 0 ? o1 : o2;
          o[ooff[mt][h] + k] = float2{acc[mt][2 * h], acc[mt][2 * h + 1]};
        }
    }
  }
}

// FP32 rotation of larger blocks on the matrix cores (v_mfma_f32_16x16x4f32),
// in the same real form out = C psi: rows M = (matrix m, state j, re/im) of
// the result, columns k = (state i, re/im) of the input.  A workgroup takes
// RB_PTS consecutive grid points and stages every input column of them in
// LDS ([k][point], all loaded before the first store, so out1 may alias
// psi); its four wavefronts take the 16-row tiles mt = w, w + 4, ... (at most
// MTW each) over the four 16-point subtiles.  Per k-step of four input
// columns a lane needs C[16 mt + l % 16][4 ks + l / 16], one FP64 entry of
// U_m (re or im, sign by the 2 x 2 block) read straight from U (L2
// resident) one k-step ahead of its products.  Result lane l, register r =
// row 16 mt + 4 (l / 16) + r at point 16 s + l % 16: registers (0, 1) and
// (2, 3) are the (re, im) of one state, stored as one 8-byte value.
constexpr int RB_PTS = 64;
constexpr int RB_MTW_MAX = 8;

template <int MTW>
__global__ void __launch_bounds__(256) rotate_mfma_big_kernel(RotArgs ra) {
  extern __shared__ float rb_lds[];  // [4 ks_n][RB_PTS]
  const RotIso& ri = ra.iso[blockIdx.y];
  const int n = ri.nst;
  const int64_t K = ra.K;
  const int64_t p0 = (int64_t)blockIdx.x * RB_PTS;
  if (n <= 0 || p0 >= K) return;  // uniform over the workgroup
  const bool two = ri.U2 != nullptr;
  const int kin = 2 * n, ks_n = (kin + 3) / 4, kpad = 4 * ks_n;
  const int mtot = (two ? 4 : 2) * n, mt_n = (mtot + 15) / 16;
  // stage the input: complex (state i, point p) -> rows 2i (re), 2i + 1 (im)
  const float2* psi = (const float2*)ri.psi;
  for (int e = threadIdx.x; e < (kpad / 2) * RB_PTS; e += 256) {
    const int i = e / RB_PTS, p = e % RB_PTS;
    float2 v = {0.f, 0.f};
    if (i < n && p0 + p < K) v = psi[(int64_t)i * K + p0 + p];
    rb_lds[(2 * i) * RB_PTS + p] = v.x;
    rb_lds[(2 * i + 1) * RB_PTS + p] = v.y;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = (lane >> 4) & 1;  // input re/im of this lane (k = 4 ks + lane / 16)
  // per tile q: this lane's row -> U_m element offset of (i = 0, j), which
  // double (re or im) and the sign; rows past mtot give zero
  const double* ub[MTW];
  double sg[MTW];
#pragma unroll
  for (int q = 0; q < MTW; ++q) {
    const int mt = w + 4 * q;
    const int M = 16 * mt + (lane & 15);
    const bool ok = mt < mt_n && M < mtot;
    const int Mc = ok ? M : 0;
    const int m = Mc / kin, jj = Mc % kin, j = jj >> 1, co = jj & 1;
    const double* U = m == 0 ? ri.U1 : ri.U2;
    ub[q] = U + 2 * j + (co == c ? 0 : 1);
    sg[q] = ok ? ((co == 0 && c == 1) ? -1.0 : 1.0) : 0.0;
  }
  rv4f acc[MTW][4];
#pragma unroll
  for (int q = 0; q < MTW; ++q)
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[q][s] = rv4f{0.f, 0.f, 0.f, 0.f};
  // input state of this lane at k-step ks: i = 2 ks + lane / 32 (clamped; a
  // padded column has a zero B row, so its A value is irrelevant)
  auto uload = [&](double (&v)[MTW], int ks) {
    const int i = min(2 * ks + (lane >> 5), n - 1);
#pragma unroll
    for (int q = 0; q < MTW; ++q) v[q] = ub[q][2 * (int64_t)i * n];
  };
  double vn[MTW];
  uload(vn, 0);
  for (int ks = 0; ks < ks_n; ++ks) {
    float a[MTW];
#pragma unroll
    for (int q = 0; q < MTW; ++q) a[q] = (float)(sg[q] * vn[q]);
    if (ks + 1 < ks_n) uload(vn, ks + 1);
    float b[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) b[s] = rb_lds[(4 * ks + (lane >> 4)) * RB_PTS + 16 * s + (lane & 15)];
#pragma unroll
    for (int q = 0; q < MTW; ++q) {
      if (w + 4 * q >= mt_n) continue;
#pragma unroll
      for (int s = 0; s < 4; ++s) acc[q][s] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q], b[s], acc[q][s], 0, 0, 0);
    }
  }
  float2* o1 = (float2*)ri.out1;
  float2* o2 = (float2*)ri.out2;
#pragma unroll
  for (int q = 0; q < MTW; ++q) {
    const int mt = w + 4 * q;
    if (mt >= mt_n) continue;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int M = 16 * mt + 4 * (lane >> 4) + 2 * h;
      if (M >= mtot) continue;
      const int m = M / kin, j = (M % kin) >> 1;
      float2* o = m == 0 ? o1 : o2;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const int64_t p = p0 + 16 * s + (lane & 15);
        if (p < K) o[(int64_t)j * K + p] = float2{acc[q][s][2 * h], acc[q][s][2 * h + 1]};
      }
    }
  }
}

// FP64 rotation on the FP64 matrix cores (v_mfma_f64_16x16x4f64), the real
// form of rotate_mfma_big_kernel: a workgroup stages every input column of
// RD_PTS consecutive grid points in LDS (FP64 rows (state i, re/im), all
// loaded before the first store, so out1 may alias psi) and its four
// wavefronts take the 16-row tiles mt = w, w + 4, ... (at most MTW each)
// over the RD_SUB 16-point subtiles.  The FP64 product leaves lane l with
// rows l / 16 + 4 r (r = 0..3) of a tile, so the coefficient rows are
// permuted: physical row p of a tile carries logical row 4 (p % 4) + p / 4,
// and lane l then holds logical rows 4 (l / 16) + r, the (re, im) of two
// output states, stored as two 16-byte values.  Per k-step of four input
// columns a lane needs one FP64 entry of U_m (re or im, sign by the 2 x 2
// block of the complex product), read from U (L2 resident) one k-step
// ahead of its products.
constexpr int RD_SUB = 2;
constexpr int RD_PTS = 16 * RD_SUB;
constexpr int RD_MTW_MAX = 8;

template <int MTW>
__global__ void __launch_bounds__(256) rotate_mfma_f64_kernel(RotArgs ra) {
  typedef double rd4 __attribute__((ext_vector_type(4)));
  extern __shared__ double rd_lds[];  // [4 ks_n][RD_PTS]
  const RotIso& ri = ra.iso[blockIdx.y];
  const int n = ri.nst;
  const int64_t K = ra.K;
  const int64_t p0 = (int64_t)blockIdx.x * RD_PTS;
  if (n <= 0 || p0 >= K) return;  // uniform over the workgroup
  const bool two = ri.U2 != nullptr;
  const int kin = 2 * n, ks_n = (kin + 3) / 4, kpad = 4 * ks_n;
  const int mtot = (two ? 4 : 2) * n, mt_n = (mtot + 15) / 16;
  // stage the input: complex (state i, point p) -> rows 2i (re), 2i + 1 (im)
  const cx<double>* psi = (const cx<double>*)ri.psi;
  for (int e = threadIdx.x; e < (kpad / 2) * RD_PTS; e += 256) {
    const int i = e / RD_PTS, p = e % RD_PTS;
    cx<double> v = {0.0, 0.0};
    if (i < n && p0 + p < K) v = psi[(int64_t)i * K + p0 + p];
    rd_lds[(2 * i) * RD_PTS + p] = v.x;
    rd_lds[(2 * i + 1) * RD_PTS + p] = v.y;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int c = (lane >> 4) & 1;  // input re/im of this lane (k = 4 ks + lane / 16)
  // per tile q: this lane's A row (physical lane % 16, logical 4 (p % 4) +
  // p / 4) -> U_m element offset of (i = 0, j), which double and the sign;
  // rows past mtot give zero
  const double* ub[MTW];
  double sg[MTW];
  const int pr = lane & 15, lr = 4 * (pr & 3) + (pr >> 2);
#pragma unroll
  for (int q = 0; q < MTW; ++q) {
    const int mt = w + 4 * q;
    const int M = 16 * mt + lr;
    const bool ok = mt < mt_n && M < mtot;
    const int Mc = ok ? M : 0;
    const int m = Mc / kin, jj = Mc % kin, j = jj >> 1, co = jj & 1;
    const double* U = m == 0 ? ri.U1 : ri.U2;
    ub[q] = U + 2 * j + (co == c ? 0 : 1);
    sg[q] = ok ? ((co == 0 && c == 1) ? -1.0 : 1.0) : 0.0;
  }
  rd4 acc[MTW][RD_SUB];
#pragma unroll
  for (int q = 0; q < MTW; ++q)
#pragma unroll
    for (int s = 0; s < RD_SUB; ++s) acc[q][s] = rd4{0.0, 0.0, 0.0, 0.0};
  // input state of this lane at k-step ks: i = 2 ks + lane / 32 (clamped; a
  // padded column has a zero B row, so its A value is irrelevant)
  auto uload = [&](double (&v)[MTW], int ks) {
    const int i = min(2 * ks + (lane >> 5), n - 1);
#pragma unroll
    for (int q = 0; q < MTW; ++q) v[q] = ub[q][2 * (int64_t)i * n];
  };
  double vn[MTW];
  uload(vn, 0);
  for (int ks = 0; ks < ks_n; ++ks) {
    double a[MTW];
#pragma unroll
    for (int q = 0; q < MTW; ++q) a[q] = sg[q] * vn[q];
    if (ks + 1 < ks_n) uload(vn, ks + 1);
    double b[RD_SUB];
#pragma unroll
    for (int s = 0; s < RD_SUB; ++s) b[s] = rd_lds[(4 * ks + (lane >> 4)) * RD_PTS + 16 * s + (lane & 15)];
#pragma unroll
    for (int q = 0; q < MTW; ++q) {
      if (w + 4 * q >= mt_n) continue;
#pragma unroll
      for (int s = 0; s < RD_SUB; ++s)
        acc[q][s] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], b[s], acc[q][s], 0, 0, 0);
    }
  }
  cx<double>* o1 = (cx<double>*)ri.out1;
  cx<double>* o2 = (cx<double>*)ri.out2;
#pragma unroll
  for (int q = 0; q < MTW; ++q) {
    const int mt = w + 4 * q;
    if (mt >= mt_n) continue;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int M = 16 * mt + 4 * (lane >> 4) + 2 * h;
      if (M >= mtot) continue;
      const int m = M / kin, j = (M % kin) >> 1;
      cx<double>* o = m == 0 ? o1 : o2;
#pragma unroll
      for (int s = 0; s < RD_SUB; ++s) {
        const int64_t p = p0 + 16 * s + (lane & 15);
        if (p < K) o[(int64_t)j * K + p] = cx<double>{acc[q][s][2 * h], acc[q][s][2 * h + 1]};
      }
    }
  }
}

}  // namespace

extern "C" int hfb_gram_workspace(int nst, long long K) {
  const int np4 = (nst + GV_T - 1) / GV_T * GV_T;
  int nc = gram_vec_nchunk(nst, K);
  if (K % 32 == 0) {  // the matrix-core path: fewest items (one Hermitian matrix) give most chunks
    const int nb = gram_mfma_blocks(nst);
    const int mc = gram_mfma_nchunk(2 * K / 64, nb * (nb + 1) / 2);
    nc = mc > nc ? mc : nc;
  }
  return nc * 4 * np4 * np4 * 2;  // doubles for up to four matrices
}

// Both isospin blocks in one launch: A[iq], B[iq*4 + m], out[iq] (nm, nst_iq,
// nst_iq) complex, part[iq] workspaces of hfb_gram_workspace(nst_iq, K).
// A block with nst 0 is skipped.
namespace {
// Chunk sums of both isospin blocks (gram_vec_reduce_kernel).
int gram_reduce(const GramVecArgs& v, int nm, const int* nst, double* const* part, double* const* out, int nchunk,
                double scale, hipStream_t st) {
  GramRedIso rr[2] = {};
  int64_t tot = 0;
  for (int iq = 0; iq < 2; ++iq) {
    rr[iq].nst = nst[iq];
    if (nst[iq] <= 0) continue;
    rr[iq].part = part[iq];
    rr[iq].out = out[iq];
    rr[iq].np = (nst[iq] + GV_T - 1) / GV_T * GV_T;
    const int64_t t = (int64_t)nm * nst[iq] * nst[iq];
    tot = t > tot ? t : tot;
  }
  int nb = (int)((tot * 64 + 255) / 256);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(gram_vec_reduce_kernel, dim3(nb < 1 ? 1 : nb, 2), dim3(256), 0, st, rr[0], rr[1], nchunk, nm,
                     v.bsame, scale);
  return launch_status();
}
}  // namespace

extern "C" int hfb_gram2_ex(int prec, const void* const* A, const void* const* B, int nm, const int* nst,
                            long long K, double scale, double* const* part, double* const* out,
                            const hfb_finish_job* job, void* stream) {
  if (!A || !B || !part || !out || !nst || nm < 1 || nm > 4 || K <= 0) return HFB_E_ARG;
  const hfb_finish_job fj = job ? *job : hfb_finish_job{};
  if (job && !finish_job_ok(fj)) return HFB_E_ARG;
  GramVecArgs v{};
  v.bsame = 0;
  for (int m = 0; m < nm; ++m) {
    bool same = true;
    for (int iq = 0; iq < 2; ++iq)
      if (nst[iq] > 0 && B[iq * 4 + m] != A[iq]) same = false;
    if (same) v.bsame |= 1 << m;
  }
  int npmax = 0;
  for (int iq = 0; iq < 2; ++iq) {
    GramIso& gi = v.iso[iq];
    gi.nst = nst[iq];
    gi.np = nst[iq] > 0 ? (nst[iq] + GV_T - 1) / GV_T * GV_T : 0;
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
  if (npmax == 0) {
    if (!job) return HFB_OK;
    hipLaunchKernelGGL(gram_vec_kernel<double>, dim3(1, 2), dim3(256), 0, (hipStream_t)stream, v, fj);
    return launch_status();
  }
  if (prec != HFB_F64 && prec != HFB_F32) return HFB_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  int nsmax = nst[0] > nst[1] ? nst[0] : nst[1];
  if (prec == HFB_F32 && K % 32 == 0 && (int64_t)nsmax * K * 8 < (1LL << 32)) {  // matrix cores (gram_mfma_kernel)
    GramMfmaArgs ma{};
    ma.iso[0] = v.iso[0];
    ma.iso[1] = v.iso[1];
    ma.nm = nm;
    ma.bsame = v.bsame;
    ma.Kr = 2 * K;
    ma.ngroup = (int)(2 * K / 64);
    int nsub = 0;
    for (int iq = 0; iq < 2; ++iq) {
      GramMfmaIso& mi = ma.mi[iq];
      mi.nst = nst[iq];
      if (nst[iq] <= 0) continue;
      mi.nt = (nst[iq] + 3) / 4;
      mi.nb = gram_mfma_blocks(nst[iq]);
      for (int m = 0; m < nm; ++m) mi.nsub += ((v.bsame >> m) & 1) ? mi.nb * (mi.nb + 1) / 2 : mi.nb * mi.nb;
      nsub += mi.nsub;
    }
    const int nc0 = gram_mfma_nchunk(ma.ngroup, nsub);
    ma.gpc = (ma.ngroup + nc0 - 1) / nc0;
    ma.nchunk = (ma.ngroup + ma.gpc - 1) / ma.gpc;
    ma.nitem = (int64_t)ma.nchunk * nsub;
    const int64_t gx = (ma.nitem + 3) / 4 + (job ? 1 : 0);
    int bt = 1;  // the largest block edge of this launch
    for (int iq = 0; iq < 2; ++iq)
      if (ma.mi[iq].nst > 0) {
        const int e = (ma.mi[iq].nt + ma.mi[iq].nb - 1) / ma.mi[iq].nb;
        bt = e > bt ? e : bt;
      }
    const dim3 grid((unsigned)gx);
    auto launch = [&](auto kern, size_t lds) {
      if (lds > 32 * 1024) allow_full_lds((const void*)kern);
      hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, ma, fj);
    };
    if (bt <= 2)
      launch(gram_mfma_kernel<2>, gram_mfma_lds<2>());
    else if (bt <= 3)
      launch(gram_mfma_kernel<3>, gram_mfma_lds<3>());
    else if (bt <= 4)
      launch(gram_mfma_kernel<4>, gram_mfma_lds<4>());
    else
      launch(gram_mfma_kernel<GM_BT>, gram_mfma_lds<GM_BT>());
    int rc = launch_status();
    if (rc != HFB_OK) return rc;
    return gram_reduce(v, nm, nst, part, out, ma.nchunk, scale, st);
  }
  {  // register-tiled path (gram_vec_kernel)
    v.nm = nm;
    v.K = K;
    v.nchunk = gram_vec_nchunk(npmax, K);
    v.kc = (K + v.nchunk - 1) / v.nchunk;
    int items = 0;
    for (int iq = 0; iq < 2; ++iq) {
      const int nt = (nst[iq] + GV_T - 1) / GV_T;
      int cnt = 0;
      for (int m = 0; m < nm; ++m) cnt += ((v.bsame >> m) & 1) ? nt * (nt + 1) / 2 : nt * nt;
      v.ntile[iq] = nst[iq] > 0 ? cnt : 0;
      items = v.ntile[iq] * v.nchunk > items ? v.ntile[iq] * v.nchunk : items;
    }
    const int gx = (items + 3) / 4 + (job ? 1 : 0);
    if (prec == HFB_F64)
      hipLaunchKernelGGL(gram_vec_kernel<double>, dim3(gx, 2), dim3(256), 0, st, v, fj);
    else if (prec == HFB_F32)
      hipLaunchKernelGGL(gram_vec_kernel<float>, dim3(gx, 2), dim3(256), 0, st, v, fj);
    else
      return HFB_E_ARG;
    int rc = launch_status();
    if (rc != HFB_OK) return rc;
    return gram_reduce(v, nm, nst, part, out, v.nchunk, scale, st);
  }
}

extern "C" int hfb_gram2(int prec, const void* const* A, const void* const* B, int nm, const int* nst, long long K,
                         double scale, double* const* part, double* const* out, void* stream) {
  return hfb_gram2_ex(prec, A, B, nm, nst, K, scale, part, out, nullptr, stream);
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

namespace {

int launch_rotate(int prec, RotArgs ra, int niso, hipStream_t st) {
  if (prec != HFB_F64 && prec != HFB_F32) return HFB_E_ARG;
  int nmax = 0, nmin = 1 << 30;
  for (int iq = 0; iq < niso; ++iq) {
    const int n = ra.iso[iq].nst;
    if (n <= 0) continue;
    nmax = n > nmax ? n : nmax;
    nmin = n < nmin ? n : nmin;
  }
  if (nmax == 0) return HFB_OK;
  if (prec == HFB_F32 && nmax <= ROT_MFMA_NMAX && ra.K % 16 == 0) {
    // matrix cores: KS k-steps over the 2 nmax input columns, MT tiles over
    // the (up to) 4 nmax out columns; ROT_BPC workgroups per CU, each
    // wavefront loading ROT_TPW consecutive tiles at a time (measured at
    // Ca40: 8 tiles, one workgroup per CU 20.3 us; 2 tiles, three 24.3 us)
    const int ks = (2 * nmax + 3) / 4;
    bool two = false;
    for (int iq = 0; iq < niso; ++iq) two = two || (ra.iso[iq].nst > 0 && ra.iso[iq].U2);
    const int mt = ((two ? 4 : 2) * nmax + 15) / 16;
    const int64_t ntile = ra.K / 16;
    constexpr int tpw = ROT_TPW;
    int64_t nb = (ntile + 4 * tpw - 1) / (4 * tpw);
    const int64_t cap = (ROT_BPC * 256 + niso - 1) / niso;
    if (nb > cap) nb = cap;
    const dim3 grid((unsigned)nb, niso);
#define HFB_ROTM(KSV, MTV)                                                                     \
  if (ks <= KSV && mt <= MTV) {                                                                \
    hipLaunchKernelGGL((rotate_mfma_kernel<KSV, MTV, tpw>), grid, dim3(256), 0, st, ra);      \
    return launch_status();                                                                    \
  }
    HFB_ROTM(4, 2)
    HFB_ROTM(6, 3)
    HFB_ROTM(8, 4)
    HFB_ROTM(10, 5)
    HFB_ROTM(12, 6)
#undef HFB_ROTM
  }
  bool two_any = false;
  for (int iq = 0; iq < niso; ++iq) two_any = two_any || (ra.iso[iq].nst > 0 && ra.iso[iq].U2);
  const int mt_max = ((two_any ? 4 : 2) * nmax + 15) / 16;
  if (prec == HFB_F32 && mt_max <= 4 * RB_MTW_MAX) {
    // matrix cores, larger blocks: RB_PTS points per workgroup, every input
    // column staged in LDS
    const int mtw = (mt_max + 3) / 4;
    const size_t lds = sizeof(float) * (size_t)(4 * ((2 * nmax + 3) / 4)) * RB_PTS;
    const dim3 grid((unsigned)((ra.K + RB_PTS - 1) / RB_PTS), niso);
#define HFB_ROTB(MW)                                                          \
  if (mtw <= MW) {                                                            \
    auto kern = rotate_mfma_big_kernel<MW>;                                   \
    if (lds > 32 * 1024) allow_full_lds((const void*)kern);                   \
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, ra);                   \
    return launch_status();                                                   \
  }
    HFB_ROTB(2)
    HFB_ROTB(4)
    HFB_ROTB(6)
    HFB_ROTB(8)
#undef HFB_ROTB
  }
  if (prec == HFB_F64 && mt_max <= 4 * RD_MTW_MAX) {
    // FP64 matrix cores: RD_PTS points per workgroup, every input column
    // staged in LDS
    const int mtw = (mt_max + 3) / 4;
    const size_t lds = sizeof(double) * (size_t)(4 * ((2 * nmax + 3) / 4)) * RD_PTS;
    const dim3 grid((unsigned)((ra.K + RD_PTS - 1) / RD_PTS), niso);
#define HFB_ROTD(MW)                                                          \
  if (mtw <= MW) {                                                            \
    auto kern = rotate_mfma_f64_kernel<MW>;                                   \
    if (lds > 32 * 1024) allow_full_lds((const void*)kern);                   \
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, ra);                   \
    return launch_status();                                                   \
  }
    HFB_ROTD(1)
    HFB_ROTD(2)
    HFB_ROTD(4)
    HFB_ROTD(8)
#undef HFB_ROTD
  }
  const size_t csz = prec == HFB_F64 ? sizeof(cx<double>) : sizeof(cx<float>);
  // two k per lane while the staged tile fits in LDS, else one
  const int kpl = csz * (size_t)nmax * 128 <= 150 * 1024 ? 2 : 1;
  const size_t lds = csz * (size_t)nmax * 64 * kpl;
  if (lds > 150 * 1024) return HFB_E_SIZE;
  const int nb = (int)((ra.K + 64 * kpl - 1) / (64 * kpl));
  // outputs per wavefront pass (rj <= every block's nst: a group never
  // reaches past the last state); the four wavefronts cover a block in one
  // pass up to 32 states and loop over groups of 8 beyond
  const int rj = nmin < 3 ? 1 : nmin <= 12 ? 3 : nmin <= 20 ? 5 : 8;
  const dim3 grid(nb, niso);
#define HFB_ROT(TT, RJV, KP)                                                     \
  {                                                                              \
    auto kern = rotate_kernel<TT, RJV, KP>;                                      \
    if (lds > 32 * 1024) allow_full_lds((const void*)kern);                      \
    hipLaunchKernelGGL(kern, grid, dim3(256), lds, st, ra);                      \
  }
#define HFB_ROTK(TT, RJV) \
  if (kpl == 2) HFB_ROT(TT, RJV, 2) else HFB_ROT(TT, RJV, 1)
#define HFB_ROTP(TT)      \
  if (rj == 1) {          \
    HFB_ROTK(TT, 1)       \
  } else if (rj == 3) {   \
    HFB_ROTK(TT, 3)       \
  } else if (rj == 5) {   \
    HFB_ROTK(TT, 5)       \
  } else {                \
    HFB_ROTK(TT, 8)       \
  }
  if (prec == HFB_F64) {
    HFB_ROTP(double)
  } else {
    HFB_ROTP(float)
  }
#undef HFB_ROTP
#undef HFB_ROTK
#undef HFB_ROT
  return launch_status();
}

}  // namespace

extern "C" int hfb_rotate(int prec, const void* psi, const double* U1, const double* U2, void* out1, void* out2,
                          int nst, long long K, void* stream) {
  if (!psi || !U1 || !out1 || nst <= 0 || K <= 0) return HFB_E_ARG;
  if (U2 && !out2) return HFB_E_ARG;
  RotArgs ra{};
  ra.iso[0] = RotIso{psi, U1, U2, out1, out2, nst};
  ra.K = K;
  return launch_rotate(prec, ra, 1, (hipStream_t)stream);
}

extern "C" int hfb_rotate2(int prec, const void* const* psi, const double* const* U1, const double* const* U2,
                           void* const* out1, void* const* out2, const int* nst, long long K, void* stream) {
  if (!psi || !U1 || !U2 || !out1 || !out2 || !nst || K <= 0) return HFB_E_ARG;
  RotArgs ra{};
  for (int iq = 0; iq < 2; ++iq) {
    if (nst[iq] < 0) return HFB_E_ARG;
    if (nst[iq] == 0) continue;
    if (!psi[iq] || !U1[iq] || !out1[iq] || (U2[iq] && !out2[iq])) return HFB_E_ARG;
    ra.iso[iq] = RotIso{psi[iq], U1[iq], U2[iq], out1[iq], out2[iq], nst[iq]};
  }
  ra.K = K;
  return launch_rotate(prec, ra, 2, (hipStream_t)stream);
}
