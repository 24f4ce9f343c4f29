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
// the result).  `scratch` needs blockDim.x/64 * nv doubles of LDS.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* scratch) {
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int j = 0; j < NV; ++j) {
    double s = v[j];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if (lane == 0) scratch[wid * NV + j] = s;
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

}  // namespace hfb
