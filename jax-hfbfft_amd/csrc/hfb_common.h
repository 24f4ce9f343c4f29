// Shared device code for the MI355X (gfx950) Skyrme-HF kernels: complex
// arithmetic, the LDS-resident mixed-radix Stockham FFT used for every
// one-axis spectral pass, and the pencil-tile geometry of the grid arrays.
//
// Layout of every spinor array: (S, 2, nx, ny, nz), complex, z fastest --
// the reference's Levels.psi layout (levels.py:62-65).  Real fields are
// (nx, ny, nz) (z fastest) or stacks of them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hfb {

template <typename T>
struct cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
// a * (i s)
template <typename T>
__device__ __forceinline__ cx<T> cmul_is(cx<T> a, T s) { return {-a.y * s, a.x * s}; }
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.x, -a.y}; }
// conj(a) * b
template <typename T>
__device__ __forceinline__ cx<T> cdotc(cx<T> a, cx<T> b) {
  return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <typename T>
__device__ __forceinline__ T cabs2(cx<T> a) { return a.x * a.x + a.y * a.y; }

// ---------------------------------------------------------------------------
// Radix plans: N = product of radices, applied in order (Stockham, DIT).
// ---------------------------------------------------------------------------
template <int N>
struct Plan;
template <> struct Plan<8>  { static constexpr int np = 2; static constexpr int r[4] = {4, 2, 0, 0}; };
template <> struct Plan<16> { static constexpr int np = 2; static constexpr int r[4] = {4, 4, 0, 0}; };
template <> struct Plan<24> { static constexpr int np = 3; static constexpr int r[4] = {4, 3, 2, 0}; };
template <> struct Plan<32> { static constexpr int np = 3; static constexpr int r[4] = {4, 4, 2, 0}; };
template <> struct Plan<48> { static constexpr int np = 3; static constexpr int r[4] = {4, 4, 3, 0}; };
template <> struct Plan<64> { static constexpr int np = 3; static constexpr int r[4] = {4, 4, 4, 0}; };
template <> struct Plan<96> { static constexpr int np = 4; static constexpr int r[4] = {4, 4, 3, 2}; };

// R-point DFT in place, X_t = sum_s u_s exp(SIGN 2 pi i s t / R).
template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft_small(cx<T>* u) {
  if constexpr (R == 2) {
    cx<T> a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
  } else if constexpr (R == 3) {
    const T c = T(-0.5);
    const T s = T(SIGN) * T(0.86602540378443864676372317075294);
    cx<T> a = u[0], b = u[1], d = u[2];
    cx<T> sp = cadd(b, d), sm = csub(b, d);
    cx<T> m = {a.x + c * sp.x, a.y + c * sp.y};
    cx<T> r = cmul_is(sm, s);  // i s (b - d)
    u[0] = cadd(a, sp);
    u[1] = cadd(m, r);
    u[2] = csub(m, r);
  } else if constexpr (R == 4) {
    cx<T> t0 = cadd(u[0], u[2]), t1 = csub(u[0], u[2]);
    cx<T> t2 = cadd(u[1], u[3]);
    cx<T> t3 = cmul_is(csub(u[1], u[3]), T(SIGN));  // (SIGN i)(u1 - u3)
    u[0] = cadd(t0, t2);
    u[2] = csub(t0, t2);
    u[1] = cadd(t1, t3);
    u[3] = csub(t1, t3);
  }
}

// One Stockham pass and the recursion over the remaining ones.  The tile is
// an LDS array [N][ldq] of complex values; columns q < Q are independent
// pencils.  tw[m] = exp(-2 pi i m / N).  Returns the buffer holding the
// result (src or dst of the first call).
template <typename T, int N, int Q, int SIGN, int PASS, int SPAN>
__device__ __forceinline__ cx<T>* fft_pass(cx<T>* src, cx<T>* dst, const cx<T>* tw, int ldq, int tid, int nthr) {
  if constexpr (PASS >= Plan<N>::np) {
    return src;
  } else {
    constexpr int R = Plan<N>::r[PASS];
    constexpr int M = N / R;
    constexpr int TWS = N / (SPAN * R);
    for (int w = tid; w < Q * M; w += nthr) {
      const int q = w % Q;
      const int i = w / Q;
      const int k = i % SPAN;
      cx<T> u[R];
#pragma unroll
      for (int r = 0; r < R; ++r) u[r] = src[(i + r * M) * ldq + q];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> t = tw[r * k * TWS];
        if constexpr (SIGN > 0) t.y = -t.y;
        u[r] = cmul(u[r], t);
      }
      dft_small<T, R, SIGN>(u);
      const int j = (i / SPAN) * SPAN * R + k;
#pragma unroll
      for (int t = 0; t < R; ++t) dst[(j + t * SPAN) * ldq + q] = u[t];
    }
    __syncthreads();
    return fft_pass<T, N, Q, SIGN, PASS + 1, SPAN * R>(dst, src, tw, ldq, tid, nthr);
  }
}

// Forward (SIGN=-1, unnormalised) or inverse (SIGN=+1, unnormalised) DFT of
// Q pencils of length N held in LDS.  Callers must __syncthreads() before;
// every thread of the block must call (the passes synchronise the block).
// (tid, nthr) select the threads that do the work, so several independent
// tiles can be transformed by sub-groups of one block.
template <typename T, int N, int Q, int SIGN>
__device__ __forceinline__ cx<T>* fft_lds(cx<T>* a, cx<T>* b, const cx<T>* tw, int ldq, int tid, int nthr) {
  return fft_pass<T, N, Q, SIGN, 0, 1>(a, b, tw, ldq, tid, nthr);
}

template <typename T, int N, int Q, int SIGN>
__device__ __forceinline__ cx<T>* fft_lds(cx<T>* a, cx<T>* b, const cx<T>* tw, int ldq) {
  return fft_pass<T, N, Q, SIGN, 0, 1>(a, b, tw, ldq, (int)threadIdx.x, (int)blockDim.x);
}

template <typename T>
__device__ __forceinline__ void load_twiddles(cx<T>* tw_lds, const cx<T>* tw_glob, int n) {
  for (int m = threadIdx.x; m < n; m += blockDim.x) tw_lds[m] = tw_glob[m];
}

// ---------------------------------------------------------------------------
// Lane-split pencil FFT (wavefront-shuffle butterflies).
//
// A pencil of length N lives in a group of L consecutive lanes; lane j of the
// group holds x[j + L m], m < N1 = N / L, in registers.  The forward DFT is
// the four-step split N = N1 * L:
//   1. an N1-point DFT inside each lane (compile-time indices: the compiler
//      keeps the pencil in VGPRs and reads twiddles through the scalar cache),
//   2. twiddle by W_N^{j k1},
//   3. an L-point DFT across the group as radix-2 decimation-in-frequency
//      butterflies between lanes j and j^h (DPP quad permutes for h = 1, 2,
//      ds_swizzle for h = 4).
// Lane j then holds X[k1 + N1 * bitrev(j)] at element k1.  The inverse runs
// the steps backwards (decimation-in-time across lanes on the bit-reversed
// input, conjugate twiddles, inverse in-lane DFT) and returns N x in the
// original layout.  Both directions are unnormalised.  tw is the table
// exp(-2 pi i m / N), m < N, in global memory.
// ---------------------------------------------------------------------------
// Wait until at most N vector memory operations of this wavefront are in
// flight (they complete in order).  The compiler tracks this wait (it adds
// none of its own for loads older than the N youngest) but not the LDS that
// global_load_lds writes: the empty asm statements keep every memory access
// on its side of the wait.
template <int N>
__device__ __forceinline__ void vm_wait() {
  static_assert(N >= 0 && N < 64, "vmcnt field");
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);  // expcnt and lgkmcnt not waited on
  asm volatile("" ::: "memory");
}

typedef __attribute__((address_space(3))) void* lds_void;

template <int H>
__device__ __forceinline__ int lane_xor_i(int v) {
  if constexpr (H == 1)
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
  else if constexpr (H == 2)
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
  else
    return __builtin_amdgcn_ds_swizzle(v, 0x1F | (H << 10));  // bit mode, xor H within 32 lanes
}

template <int H>
__device__ __forceinline__ float lane_xor(float v) {
  return __int_as_float(lane_xor_i<H>(__float_as_int(v)));
}

template <int H>
__device__ __forceinline__ double lane_xor(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_xor_i<H>((int)(b & 0xffffffffLL));
  const int hi = lane_xor_i<H>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int H, typename T>
__device__ __forceinline__ cx<T> lane_xor(cx<T> v) {
  return {lane_xor<H>(v.x), lane_xor<H>(v.y)};
}

// Sum over the L lanes of a group; every lane receives the sum.
template <int L, typename T>
__device__ __forceinline__ T group_sum(T v) {
  if constexpr (L >= 2) v += lane_xor<1>(v);
  if constexpr (L >= 4) v += lane_xor<2>(v);
  if constexpr (L >= 8) v += lane_xor<4>(v);
  return v;
}

// In-lane DFT of the NC values v[0..NC), X_t = sum_s v_s exp(SIGN 2 pi i s t / NC),
// NC | N, by recursive decimation in time (radix 4, 2 or 3).
template <typename T, int NC, int SIGN, int N>
__device__ __forceinline__ void dft_lane(cx<T>* v, const cx<T>* __restrict__ tw) {
  if constexpr (NC == 2 || NC == 3 || NC == 4) {
    dft_small<T, NC, SIGN>(v);
  } else if constexpr (NC > 4) {
    constexpr int R = (NC % 4 == 0) ? 4 : ((NC % 2 == 0) ? 2 : 3);
    constexpr int M = NC / R;
    cx<T> sub[R][M];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int m = 0; m < M; ++m) sub[r][m] = v[r + R * m];
#pragma unroll
    for (int r = 0; r < R; ++r) dft_lane<T, M, SIGN, N>(sub[r], tw);
#pragma unroll
    for (int k1 = 0; k1 < M; ++k1) {
      cx<T> u[R];
      u[0] = sub[0][k1];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        if (k1 == 0) {
          u[r] = sub[r][0];
        } else {
          cx<T> t = tw[(r * k1 * (N / NC)) % N];
          if constexpr (SIGN > 0) t.y = -t.y;
          u[r] = cmul(sub[r][k1], t);
        }
      }
      dft_small<T, R, SIGN>(u);
#pragma unroll
      for (int t = 0; t < R; ++t) v[k1 + M * t] = u[t];
    }
  }
}

template <typename T, int N, int L>
struct LaneFFT {
  static constexpr int N1 = N / L;
  static_assert(N % L == 0 && (L == 1 || L == 2 || L == 4 || L == 8), "lane split");
  static constexpr int NS = L == 8 ? 3 : (L == 4 ? 2 : (L == 2 ? 1 : 0));
  int j;           // position in the lane group
  int brev;        // bit-reversed j (log2 L bits)
  cx<T> twj[N1];   // W_N^{j k1}
  cx<T> tws[NS + 1];  // tws[s] = W_{2h}^{j mod h}, h = 2^s

  __device__ __forceinline__ void init(int jj, const cx<T>* __restrict__ tw) {
    j = jj;
    brev = 0;
#pragma unroll
    for (int b = 0; b < NS; ++b) brev |= ((jj >> b) & 1) << (NS - 1 - b);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) twj[k1] = tw[(jj * k1) % N];
#pragma unroll
    for (int st = 0; st < NS; ++st) {
      const int h = 1 << st;
      tws[st] = tw[(jj % h) * (N / (2 * h))];
    }
  }

  // frequency index of element k1 after fwd
  __device__ __forceinline__ int freq(int k1) const { return k1 + N1 * brev; }

  // Cross-lane radix-2 butterflies.  The twiddle of stage h is
  // W_{2h}^{j mod h} on the bottom lanes: 1 for h = 1, 1 or -i for h = 2
  // (applied exactly, as a swap and a sign), a table value for h = 4.  Top
  // and bottom outputs share one form, s v + u with s = -1 on the bottom.
  template <int ST>
  __device__ __forceinline__ void fwd_stages(cx<T>* v) const {
    if constexpr (ST >= 0) {
      constexpr int H = 1 << ST;
      const bool bottom = (j & H) != 0;
      const T sg = bottom ? T(-1) : T(1);
      const bool rot = bottom && (j & 1);  // times -i (stage h = 2)
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) {
        const cx<T> u = lane_xor<H>(v[k1]);
        cx<T> d = {fma(sg, v[k1].x, u.x), fma(sg, v[k1].y, u.y)};
        if constexpr (H == 2) {
          d = {rot ? d.y : d.x, rot ? -d.x : d.y};
        } else if constexpr (H >= 4) {
          if (bottom) d = cmul(d, tws[ST]);
        }
        v[k1] = d;
      }
      fwd_stages<ST - 1>(v);
    }
  }

  template <int ST>
  __device__ __forceinline__ void inv_stages(cx<T>* v) const {
    if constexpr (ST < NS) {
      constexpr int H = 1 << ST;
      const bool bottom = (j & H) != 0;
      const T sg = bottom ? T(-1) : T(1);
      const bool rot = bottom && (j & 1);  // times +i (stage h = 2)
      const cx<T> wc = cconj(tws[ST]);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) {
        cx<T> sv = v[k1];
        if constexpr (H == 2) {
          sv = {rot ? -sv.y : sv.x, rot ? sv.x : sv.y};
        } else if constexpr (H >= 4) {
          if (bottom) sv = cmul(sv, wc);
        }
        const cx<T> r = lane_xor<H>(sv);
        v[k1] = {fma(sg, sv.x, r.x), fma(sg, sv.y, r.y)};
      }
      inv_stages<ST + 1>(v);
    }
  }

  __device__ __forceinline__ void fwd(cx<T>* v, const cx<T>* __restrict__ tw) const {
    dft_lane<T, N1, -1, N>(v, tw);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], twj[k1]);
    fwd_stages<NS - 1>(v);
  }

  __device__ __forceinline__ void inv(cx<T>* v, const cx<T>* __restrict__ tw) const {
    inv_stages<0>(v);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) v[k1] = cmul(v[k1], cconj(twj[k1]));
    dft_lane<T, N1, +1, N>(v, tw);
  }
};

// ---------------------------------------------------------------------------
// Pencil tiles.  For axis a a tile holds P pencils of length N at a fixed
// value of one transverse coordinate and P consecutive values of the other:
//   a = x: tile (y, z0..z0+P)      element (k,p) at k*ny*nz + y*nz + z0+p
//   a = y: tile (x, z0..z0+P)      element (k,p) at x*ny*nz + k*nz + z0+p
//   a = z: tile (x, y0..y0+P)      element (k,p) at x*ny*nz + (y0+p)*nz + k
// ---------------------------------------------------------------------------
struct Tile {
  int64_t base;  // offset of element (0,0) inside one (nx,ny,nz) field
  int64_t sk;    // stride along the pencil
  int64_t sp;    // stride between neighbouring pencils
};

template <int AXIS>
__device__ __forceinline__ Tile make_tile(int tile, int nx, int ny, int nz, int P) {
  Tile t;
  if constexpr (AXIS == 0) {
    const int nzt = nz / P;
    const int y = tile / nzt, z0 = (tile % nzt) * P;
    t.base = (int64_t)y * nz + z0;
    t.sk = (int64_t)ny * nz;
    t.sp = 1;
  } else if constexpr (AXIS == 1) {
    const int nzt = nz / P;
    const int x = tile / nzt, z0 = (tile % nzt) * P;
    t.base = (int64_t)x * ny * nz + z0;
    t.sk = nz;
    t.sp = 1;
  } else {
    const int nyt = ny / P;
    const int x = tile / nyt, y0 = (tile % nyt) * P;
    t.base = (int64_t)x * ny * nz + (int64_t)y0 * nz;
    t.sk = 1;
    t.sp = nz;
  }
  return t;
}

template <int AXIS>
__host__ __device__ __forceinline__ int tiles_per_field(int nx, int ny, int nz, int P) {
  if (AXIS == 0) return ny * (nz / P);
  if (AXIS == 1) return nx * (nz / P);
  return nx * (ny / P);
}

// Element order for global<->LDS moves: consecutive threads walk the
// contiguous direction in memory (p for axes x,y; k for axis z).
template <int AXIS, int N, int P>
__device__ __forceinline__ void elem_of(int e, int& k, int& p) {
  if constexpr (AXIS == 2) {
    k = e % N;
    p = e / N;
  } else {
    p = e % P;
    k = e / P;
  }
}

// Coordinates (x,y,z) of tile element (k,p).
template <int AXIS>
__device__ __forceinline__ void coords_of(int tile, int k, int p, int nx, int ny, int nz, int P, int& x,
                                          int& y, int& z) {
  if constexpr (AXIS == 0) {
    const int nzt = nz / P;
    x = k;
    y = tile / nzt;
    z = (tile % nzt) * P + p;
  } else if constexpr (AXIS == 1) {
    const int nzt = nz / P;
    x = tile / nzt;
    y = k;
    z = (tile % nzt) * P + p;
  } else {
    const int nyt = ny / P;
    x = tile / nyt;
    y = (tile % nyt) * P + p;
    z = k;
  }
}

// First-derivative multiplier i*m*kfac/n with m in (-n/2, n/2); zero at m=0
// and at the Nyquist index (levels.py:142-146).  Returns the real factor f
// such that the multiplier is i*f.
template <typename T, int N>
__device__ __forceinline__ T d1_factor(int k, T kfac) {
  if (k == 0 || k == N / 2) return T(0);
  const int m = k < N / 2 ? k : k - N;
  return T(m) * kfac / T(N);
}

// k^2 with the Nyquist mode kept (levels.py:407-420), positive.
template <typename T, int N>
__device__ __forceinline__ T k2_factor(int k, T kfac) {
  const int m = k <= N / 2 ? k : N - k;
  return (T(m) * kfac) * (T(m) * kfac);
}

// Block-wide sum of `nv` values per thread into out[0..nv) (thread 0 holds
// the result).  `scratch` needs blockDim.x/64 * nv doubles of LDS.  The
// wavefront sums run level by level over all values (xor butterflies: DPP
// within quads, swizzles within 32 lanes, a permute across the halves), so
// the nv chains overlap; every lane ends with its wavefront's sum.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* scratch) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] += lane_xor<1>(v[j]);
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] += lane_xor<2>(v[j]);
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] += lane_xor<4>(v[j]);
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] += lane_xor<8>(v[j]);
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] += lane_xor<16>(v[j]);
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] += __shfl_xor(v[j], 32, 64);
  if (lane == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) scratch[wid * NV + j] = v[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += scratch[w * NV + j];
      v[j] = s;
    }
  }
  __syncthreads();
}

// Single-pass reductions across workgroups.  A producer
// that has written its share calls *_arrive on a counter with the number of
// producers; the call returns true in exactly one of them, the last to
// arrive, which may then read every share written before the other
// arrivals.  The last arrival resets the counter to zero, so every launch
// (and every replay of a captured launch) starts from a clean counter.
// Release: each thread's device-scope fence before the count; acquire: the
// fence in the last producer after it.
__device__ __forceinline__ bool block_arrive(unsigned int* counter, unsigned int expected) {
  __shared__ unsigned int arrive_last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int prev = atomicAdd(counter, 1u);
    const bool last = prev + 1u == expected;
    if (last) atomicExch(counter, 0u);
    arrive_last = last ? 1u : 0u;
  }
  __syncthreads();
  const bool last = arrive_last != 0u;
  if (last) __threadfence();
  return last;
}

}  // namespace hfb
