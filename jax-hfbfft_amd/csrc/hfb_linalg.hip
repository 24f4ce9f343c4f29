// Small dense FP64 linear algebra of diagstep, one workgroup per isospin
// (static.py:261-577): Loewdin S^{-1/2} with the eigenvalue clamp, the
// optional subspace diagonalisation of lambda_lin, U^H H_old U, the gap
// matrix, lambda = H diag(w) - G diag(wuv) split into its Hermitian and
// anti-Hermitian parts, the fluctuation measures, and the matrix M2 = U
// lambda+ that turns psi_old into the Lagrange field.
//
// Hermitian eigenproblems use a parallel cyclic Jacobi method (round-robin
// pair ordering, complex rotations) on the Hermitian matrix built from the
// lower triangle, like LAPACK's heev with uplo='L'.  Eigenvalues are
// returned in ascending order with matching eigenvector columns.
#include <math.h>

#include "hfb_ctx.h"

using namespace hfb;

namespace {

constexpr int NMAXJ = 256;

// Jacobi stops when the squared off-diagonal Frobenius norm is below
// JACOBI_TOL2 times the squared diagonal one (relative off-diagonal norm
// 1e-13): the entries of the subspace matrices carry larger errors than that
// (FP64 sums of K = 2N products, about sqrt(K) eps = 6e-14 relative at 32^3,
// and FP32 inputs on the mixed path), so further sweeps only polish
// round-off.  Measured at Ca40: two sweeps reach 1e-28 .. 1e-26 in the
// steady iterations, the third went on to 1e-36.
constexpr double JACOBI_TOL2 = 1e-26;

struct DiagIso {
  int nst, start;
  const cx<double>* S;
  const cx<double>* Hold;
  const cx<double>* Lam;   // null unless diagonalising
  const cx<double>* Gold;  // null unless pairing
  cx<double>* U;
  cx<double>* M2;
  cx<double>* W;  // 5 n x n work matrices in global memory (blocks above NDIAG_LDS states)
  cx<double>* Hm;  // static arrays of this isospin, leading dimension nmax
  cx<double>* Gm;
  cx<double>* Sym;
  cx<double>* Ls;
};

struct DiagArgs {
  DiagIso iso[2];
  int nmax;
  const double* wocc;
  const double* wstates;
  const double* wguv;
  const double* pairwg;
  double* sp_norm;
  double* sp_energy;
  double* deltaf;
  double* efl1q;
  double* efl2q;
  double* efl1;
  double* efl2;
  unsigned int* sync;  // arrival counter of the two isospin blocks
  double* stats;  // optional: [Jacobi sweeps (2), Newton-Schulz iterations (2, -1 = fell back)]
  int tdiag, ipair, tbcs, construct;
  int mode;  // DIAG_FULL, DIAG_LOEWDIN (S -> U only) or DIAG_GIVEN_U (construct from U in d.U)
};

enum { DIAG_FULL = 0, DIAG_LOEWDIN = 1, DIAG_GIVEN_U = 2 };


// Work split of an n x n matrix (n <= NP) over the NT threads: thread t
// owns column j = t % NP and rows t / NP + RG*s.  Matrices are stored densely
// with leading dimension n.
#define HFB_TILE_LOOP(NP, NT, n, i, j)                               \
  for (int j = threadIdx.x % (NP), i = threadIdx.x / (NP); j < (n) && i < (n); i += (NT) / (NP))

// 1/sqrt(x) and 1/x for x > 0: the hardware estimates refined by two
// Newton steps to double precision (a few ulp), far shorter chains than
// the correctly rounded library sqrt and division.
__device__ __forceinline__ double rsq_step(double x, double y) { return y * fma(-0.5 * x * y, y, 1.5); }
__device__ __forceinline__ double rsq_full(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  return rsq_step(x, rsq_step(x, y));
}
__device__ __forceinline__ double rcp_full(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = y * fma(-x, y, 2.0);
  return y * fma(-x, y, 2.0);
}

template <int NP, int NT, typename F>
__device__ __forceinline__ void tile_for(int n, F&& f) {
  constexpr int RG = NT / NP;
  const int j = threadIdx.x % NP;
  if (j >= n) return;
  for (int i = threadIdx.x / NP; i < n; i += RG) f(i, j);
}

// Sum of two values over the block, returned to every thread.
template <int NT>
__device__ __forceinline__ void block_sum2_all(double& a, double& b, double* red) {
  a += lane_xor<1>(a);
  b += lane_xor<1>(b);
  a += lane_xor<2>(a);
  b += lane_xor<2>(b);
  a += lane_xor<4>(a);
  b += lane_xor<4>(b);
  a += lane_xor<8>(a);
  b += lane_xor<8>(b);
  a += lane_xor<16>(a);
  b += lane_xor<16>(b);
  a += __shfl_xor(a, 32, 64);
  b += __shfl_xor(b, 32, 64);
  const int wid = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[2 * wid] = a;
    red[2 * wid + 1] = b;
  }
  __syncthreads();
  a = b = 0.0;
  for (int w = 0; w < NT / 64; ++w) {
    a += red[2 * w];
    b += red[2 * w + 1];
  }
  __syncthreads();
}

template <int NT>
__device__ __forceinline__ double block_max_all(double v, double* red) {
  v = fmax(v, lane_xor<1>(v));
  v = fmax(v, lane_xor<2>(v));
  v = fmax(v, lane_xor<4>(v));
  v = fmax(v, lane_xor<8>(v));
  v = fmax(v, lane_xor<16>(v));
  v = fmax(v, __shfl_xor(v, 32, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  v = red[0];
  for (int w = 1; w < NT / 64; ++w) v = fmax(v, red[w]);
  __syncthreads();
  return v;
}

// A (n x n, LDS) <- Hermitian matrix from the lower triangle of M (global).
template <int NP, int NT>
__device__ void load_lower_herm(cx<double>* A, const cx<double>* M, int n) {
  tile_for<NP, NT>(n, [&](int i, int j) {
    cx<double> v;
    if (i > j)
      v = M[i * n + j];
    else if (i < j)
      v = cconj(M[j * n + i]);
    else
      v = {M[i * n + i].x, 0.0};
    A[i * n + j] = v;
  });
  __syncthreads();
}

// C = op(A) B for n x n matrices (C distinct from A and B); op(A) = A or A^H.
// Each thread owns a 2x2 block of C and keeps it in registers; per k it
// reads two entries of op(A) and two of B (four complex loads for four
// complex multiply-adds, which keeps the LDS data path from being the
// limit), operands of MMK consecutive k at a time.  Odd n: the last block
// row/column reads row/column n-1 twice and stores only what exists.
__device__ __forceinline__ void mac(cx<double>& c, const cx<double> a, const cx<double> b) {
  c.x = fma(a.x, b.x, fma(-a.y, b.y, c.x));
  c.y = fma(a.x, b.y, fma(a.y, b.x, c.y));
}
template <bool CONJ_A>
__device__ __forceinline__ cx<double> mm_opa(const cx<double>* A, int n, int i, int k) {
  return CONJ_A ? cconj(A[k * n + i]) : A[i * n + k];
}

template <int NP, int NT, bool CONJ_A>
__device__ __forceinline__ void mm(cx<double>* __restrict__ C, const cx<double>* A, const cx<double>* B, int n) {
  constexpr int MMK = NT >= 1024 ? 1 : 4;  // (1024-thread blocks have 128 registers)
  const int nh = (n + 1) >> 1;
  for (int t = threadIdx.x; t < nh * nh; t += NT) {
    const int i0 = 2 * (t / nh), j0 = 2 * (t % nh);
    const int i1 = min(i0 + 1, n - 1), j1 = min(j0 + 1, n - 1);
    cx<double> c00 = {0.0, 0.0}, c01 = {0.0, 0.0}, c10 = {0.0, 0.0}, c11 = {0.0, 0.0};
    int k = 0;
    for (; k + MMK <= n; k += MMK) {
      cx<double> a0[MMK], a1[MMK], b0[MMK], b1[MMK];
#pragma unroll
      for (int c = 0; c < MMK; ++c) {
        a0[c] = mm_opa<CONJ_A>(A, n, i0, k + c);
        a1[c] = mm_opa<CONJ_A>(A, n, i1, k + c);
        b0[c] = B[(k + c) * n + j0];
        b1[c] = B[(k + c) * n + j1];
      }
#pragma unroll
      for (int c = 0; c < MMK; ++c) {
        mac(c00, a0[c], b0[c]);
        mac(c01, a0[c], b1[c]);
        mac(c10, a1[c], b0[c]);
        mac(c11, a1[c], b1[c]);
      }
    }
    for (; k < n; ++k) {
      const cx<double> a0 = mm_opa<CONJ_A>(A, n, i0, k), a1 = mm_opa<CONJ_A>(A, n, i1, k);
      const cx<double> b0 = B[k * n + j0], b1 = B[k * n + j1];
      mac(c00, a0, b0);
      mac(c01, a0, b1);
      mac(c10, a1, b0);
      mac(c11, a1, b1);
    }
    C[i0 * n + j0] = c00;
    if (j0 + 1 < n) C[i0 * n + j0 + 1] = c01;
    if (i0 + 1 < n) {
      C[(i0 + 1) * n + j0] = c10;
      if (j0 + 1 < n) C[(i0 + 1) * n + j0 + 1] = c11;
    }
  }
  __syncthreads();
}

struct JacobiShared {
  // rotations of the current round
  int p[NMAXJ / 2], q[NMAXJ / 2];
  double c[NMAXJ / 2], s[NMAXJ / 2];
  cx<double> e[NMAXJ / 2];
  double red[32];
  int perm[NMAXJ];
  double ev[NMAXJ];
};

// Rotations of two consecutive rounds (buffer t & 1) for jacobi_eigh.
struct JacobiShared2 {
  int p[2][NMAXJ / 2], q[2][NMAXJ / 2];
  double c[2][NMAXJ / 2], s[2][NMAXJ / 2];
  cx<double> e[2][NMAXJ / 2];
  double red[32];
  int perm[NMAXJ];
  double ev[NMAXJ];
};

// Eigen-decomposition of the Hermitian A (n x n, dense; destroyed) for
// blocks whose work matrices live in LDS; V gets the
// eigenvectors as columns, js.ev the eigenvalues; sorted ascending when
// `sort` (T is scratch for the column permutation).  Parallel cyclic Jacobi:
// in each round the n/2 disjoint pairs (p_k, q_k) of a round-robin ordering
// are rotated together, A <- R^H A R with R block-diagonal over the pairs.
// A round computes the rotations (one thread per pair), then every 2x2
// block (pair a, pair b) of the new matrix, R_a^H A[a,b] R_b, into the
// second buffer A2 and V <- V R in place; the buffers then swap.  For odd n
// one index per round has no partner and is carried unrotated.
template <int NP, int NT>
__device__ __forceinline__ int jacobi_eigh(cx<double>* A, cx<double>* A2, cx<double>* V, int n, JacobiShared2& js,
                                           bool sort, cx<double>* T) {
  const int tid = threadIdx.x;
  tile_for<NP, NT>(n, [&](int i, int j) { V[i * n + j] = cx<double>{i == j ? 1.0 : 0.0, 0.0}; });
  const int m = n + (n & 1);
  const int npair = m / 2;
  __syncthreads();
  int sweep = 0;
  for (; sweep < 40; ++sweep) {
    double off = 0.0, dia = 0.0;
    HFB_TILE_LOOP(NP, NT, n, i, j) {
      if (i == j)
        dia += cabs2(A[i * n + j]);
      else
        off += cabs2(A[i * n + j]);
    }
    block_sum2_all<NT>(off, dia, js.red);
    if (off <= JACOBI_TOL2 * dia || off == 0.0) break;
    // rotation angles of round t into buffer b (threads < npair)
    auto rotations = [&](int t, int b) {
      if (tid < npair) {
        // round-robin pairs: index 0 fixed, the others rotate by t
        const int k = tid;
        const int m1 = m - 1;
        int xa = k - 1 + t;  // < 2 (m - 1)
        xa = xa >= m1 ? xa - m1 : xa;
        int xb = m1 - k - 1 + t;
        xb = xb >= m1 ? xb - m1 : xb;
        const int a0 = (k == 0) ? 0 : xa + 1;
        const int b0 = xb + 1;
        int p = a0 < b0 ? a0 : b0, q = a0 < b0 ? b0 : a0;
        double c = 1.0, sn = 0.0;
        cx<double> ea = {1.0, 0.0};
        if (q < n) {
          const double app = A[p * n + p].x, aqq = A[q * n + q].x;
          const cx<double> apq = A[p * n + q];
          const double r2 = cabs2(apq), den = fabs(app) + fabs(aqq);
          if (r2 > 0.0 && r2 > 1e-36 * den * den) {
            const double ri = rsq_full(r2);  // 1 / |a_pq|
            ea = {apq.x * ri, apq.y * ri};
            const double th = 0.5 * (aqq - app) * ri;
            const double h = 1.0 + th * th;
            const double rt = h * rsq_full(h);  // sqrt(1 + th^2)
            const double tt = (th >= 0.0 ? 1.0 : -1.0) * rcp_full(fabs(th) + rt);
            c = rsq_full(1.0 + tt * tt);
            sn = tt * c;
          }
        } else {
          q = -1;  // p has no partner this round (odd n)
        }
        js.p[b][k] = p;
        js.q[b][k] = q;
        js.c[b][k] = c;
        js.s[b][k] = sn;
        js.e[b][k] = ea;
      }
    };
    rotations(0, 0);
    __syncthreads();
    for (int t = 0; t < m - 1; ++t) {
      const int b = t & 1;
      // every 2x2 block (pair ka, pair kb) of R^H A R into A2
      const int nblk = npair * npair;
      for (int it = tid; it < nblk; it += NT) {
        const int ka = it / npair, kb = it % npair;
        const int pa = js.p[b][ka], qa = js.q[b][ka], pb = js.p[b][kb], qb = js.q[b][kb];
        const double ca = js.c[b][ka], sa = js.s[b][ka], cb = js.c[b][kb], sb = js.s[b][kb];
        const cx<double> eca = cscale(js.e[b][ka], ca), esa = cscale(js.e[b][ka], sa);
        const cx<double> ecb = cscale(js.e[b][kb], cb), esb = cscale(js.e[b][kb], sb);
        const cx<double> z = {0.0, 0.0};
        // rows pa, qa of A, columns pb, qb (absent partners read as zero)
        const cx<double> x00 = A[pa * n + pb];
        const cx<double> x01 = qb >= 0 ? A[pa * n + qb] : z;
        const cx<double> x10 = qa >= 0 ? A[qa * n + pb] : z;
        const cx<double> x11 = qa >= 0 && qb >= 0 ? A[qa * n + qb] : z;
        // columns: [xp xq] R_b
        const cx<double> y00 = csub(cmul(x00, ecb), cscale(x01, sb));
        const cx<double> y01 = cadd(cmul(x00, esb), cscale(x01, cb));
        const cx<double> y10 = csub(cmul(x10, ecb), cscale(x11, sb));
        const cx<double> y11 = cadd(cmul(x10, esb), cscale(x11, cb));
        // rows: R_a^H
        cx<double> n00 = csub(cmul(cconj(eca), y00), cscale(y10, sa));
        cx<double> n01 = csub(cmul(cconj(eca), y01), cscale(y11, sa));
        cx<double> n10 = cadd(cmul(cconj(esa), y00), cscale(y10, ca));
        cx<double> n11 = cadd(cmul(cconj(esa), y01), cscale(y11, ca));
        if (ka == kb && qa >= 0 && sa != 0.0) {  // the rotated pair: diagonal 2x2 block
          n01 = z;
          n10 = z;
          n00.y = 0.0;
          n11.y = 0.0;
        }
        A2[pa * n + pb] = n00;
        if (qb >= 0) A2[pa * n + qb] = n01;
        if (qa >= 0) A2[qa * n + pb] = n10;
        if (qa >= 0 && qb >= 0) A2[qa * n + qb] = n11;
      }
      __syncthreads();
      cx<double>* sw = A;
      A = A2;
      A2 = sw;
      // the next round's angles (first wavefront) beside V <- V R of this
      // round (the other wavefronts): different buffers, one barrier
      if (t + 1 < m - 1) rotations(t + 1, b ^ 1);
      for (int it = tid - 64; it >= 0 && it < n * npair; it += NT - 64) {
        const int r = it / npair, k = it % npair;
        const int p = js.p[b][k], q = js.q[b][k];
        if (q >= 0 && js.s[b][k] != 0.0) {
          const cx<double> xp = V[r * n + p], xq = V[r * n + q];
          const cx<double> ec = cscale(js.e[b][k], js.c[b][k]), es = cscale(js.e[b][k], js.s[b][k]);
          V[r * n + p] = csub(cmul(xp, ec), cscale(xq, js.s[b][k]));
          V[r * n + q] = cadd(cmul(xp, es), cscale(xq, js.c[b][k]));
        }
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < n; i += NT) js.ev[i] = A[i * n + i].x;
  __syncthreads();
  if (sort) {
    // stable ascending order by rank: index i goes to the number of
    // eigenvalues below it, ties kept in index order (the identity first, so
    // that every entry stays an index below n whatever the values)
    for (int i = tid; i < n; i += NT) js.perm[i] = i;
    __syncthreads();
    for (int i = tid; i < n; i += NT) {
      const double e = js.ev[i];
      int r = 0;
      for (int j = 0; j < n; ++j) {
        const double f = js.ev[j];
        r += (f < e || (f == e && j < i)) ? 1 : 0;
      }
      js.perm[r] = i;
    }
    __syncthreads();
    tile_for<NP, NT>(n, [&](int i, int j) { T[i * n + j] = V[i * n + js.perm[j]]; });
    __syncthreads();
    tile_for<NP, NT>(n, [&](int i, int j) { V[i * n + j] = T[i * n + j]; });
    double evs = 0.0;
    if (tid < n) evs = js.ev[js.perm[tid]];
    __syncthreads();
    if (tid < n) js.ev[tid] = evs;
    __syncthreads();
  }
  return sweep;
}

// Packed lower triangle of a Hermitian n x n matrix: entry (i, j), i >= j,
// at i (i + 1) / 2 + j; an upper entry is the conjugate of its mirror.
__device__ __forceinline__ int ploff(int i) { return (i * (i + 1)) >> 1; }
__device__ __forceinline__ cx<double> pld(const cx<double>* A, int i, int j) {
  return i >= j ? A[ploff(i) + j] : cconj(A[ploff(j) + i]);
}
__device__ __forceinline__ void pst(cx<double>* A, int i, int j, cx<double> v) {
  if (i >= j)
    A[ploff(i) + j] = v;
  else
    A[ploff(j) + i] = cconj(v);
}
// Row of linear index e of the packed triangle: the largest i with
// ploff(i) <= e (a floating-point estimate, then corrected).
__device__ __forceinline__ int prow(int e) {
  int i = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
  if (ploff(i + 1) <= e) ++i;
  if (ploff(i) > e) --i;
  return i;
}

// A (packed, n (n + 1) / 2) <- the Hermitian matrix of M's lower triangle
// (M dense, leading dimension n), diagonal made real.
template <int NT>
__device__ void load_lower_packed(cx<double>* A, const cx<double>* M, int n) {
  const int tot = ploff(n);
  for (int e = threadIdx.x; e < tot; e += NT) {
    const int i = prow(e), j = e - ploff(i);
    const cx<double> v = M[i * n + j];
    A[e] = i == j ? cx<double>{v.x, 0.0} : v;
  }
  __syncthreads();
}

// Eigen-decomposition of the Hermitian matrix in A (packed lower triangle,
// destroyed): Vout (n x n, leading dimension n) gets the eigenvectors as
// columns, js.ev the eigenvalues, in ascending order when `sort`.  Vt (n x n)
// is scratch holding V transposed (row c = eigenvector c) while the
// rotations are applied, so that the update of a column pair is a
// contiguous row pair.  Parallel cyclic Jacobi: in each round the n/2
// disjoint pairs (p_k, q_k) of a round-robin ordering are rotated together,
// A <- R^H A R with R block-diagonal over the pairs.  A round computes the
// rotations (one thread per pair), then every 2x2 block (pair a, pair b),
// a >= b, of R_a^H A[a,b] R_b in place (the blocks of one round are
// disjoint, and the packed triangle holds each entry once) beside
// V <- V R.  For odd n one index per round has no partner and is carried
// unrotated.
template <int NT>
__device__ __forceinline__ int jacobi_packed(cx<double>* __restrict__ A, cx<double>* __restrict__ Vt,
                                             cx<double>* __restrict__ Vout, int n, JacobiShared& js, bool sort) {
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += NT) {
    const int r = e / n, c = e - r * n;
    Vt[e] = cx<double>{r == c ? 1.0 : 0.0, 0.0};
  }
  const int m = n + (n & 1);
  const int npair = m / 2;
  const int nblk = ploff(npair);
  const int nv = n * npair;
  __syncthreads();
  int sweep = 0;
  for (; sweep < 40; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int e = tid; e < ploff(n); e += NT) {
      const int i = prow(e), j = e - ploff(i);
      const double v = cabs2(A[e]);
      if (i == j)
        dia += v;
      else
        off += 2.0 * v;
    }
    block_sum2_all<NT>(off, dia, js.red);
    if (off <= JACOBI_TOL2 * dia || off == 0.0) break;
    for (int t = 0; t < m - 1; ++t) {
      // rotation angles of round t (threads < npair)
      if (tid < npair) {
        // round-robin pairs: index 0 fixed, the others rotate by t
        const int k = tid;
        const int m1 = m - 1;
        int xa = k - 1 + t;  // < 2 (m - 1)
        xa = xa >= m1 ? xa - m1 : xa;
        int xb = m1 - k - 1 + t;
        xb = xb >= m1 ? xb - m1 : xb;
        const int a0 = (k == 0) ? 0 : xa + 1;
        const int b0 = xb + 1;
        int p = a0 < b0 ? a0 : b0, q = a0 < b0 ? b0 : a0;
        double c = 1.0, sn = 0.0;
        cx<double> ea = {1.0, 0.0};
        if (q < n) {
          const double app = A[ploff(p) + p].x, aqq = A[ploff(q) + q].x;
          const cx<double> apq = cconj(A[ploff(q) + p]);
          const double r2 = cabs2(apq), den = fabs(app) + fabs(aqq);
          if (r2 > 0.0 && r2 > 1e-36 * den * den) {
            const double ri = rsq_full(r2);  // 1 / |a_pq|
            ea = {apq.x * ri, apq.y * ri};
            const double th = 0.5 * (aqq - app) * ri;
            const double h = 1.0 + th * th;
            const double rt = h * rsq_full(h);  // sqrt(1 + th^2)
            const double tt = (th >= 0.0 ? 1.0 : -1.0) * rcp_full(fabs(th) + rt);
            c = rsq_full(1.0 + tt * tt);
            sn = tt * c;
          }
        } else {
          q = -1;  // p has no partner this round (odd n)
        }
        js.p[k] = p;
        js.q[k] = q;
        js.c[k] = c;
        js.s[k] = sn;
        js.e[k] = ea;
      }
      __syncthreads();
      // the blocks (ka, kb), ka >= kb, of R^H A R, in place
      for (int it = tid; it < nblk; it += NT) {
        const int ka = prow(it), kb = it - ploff(ka);
        const int pa = js.p[ka], qa = js.q[ka], pb = js.p[kb], qb = js.q[kb];
        const double ca = js.c[ka], sa = js.s[ka], cb = js.c[kb], sb = js.s[kb];
        const cx<double> eca = cscale(js.e[ka], ca), esa = cscale(js.e[ka], sa);
        const cx<double> ecb = cscale(js.e[kb], cb), esb = cscale(js.e[kb], sb);
        const cx<double> z = {0.0, 0.0};
        // rows pa, qa of A, columns pb, qb (absent partners read as zero)
        const cx<double> x00 = pld(A, pa, pb);
        const cx<double> x01 = qb >= 0 ? pld(A, pa, qb) : z;
        const cx<double> x10 = qa >= 0 ? pld(A, qa, pb) : z;
        const cx<double> x11 = qa >= 0 && qb >= 0 ? pld(A, qa, qb) : z;
        // columns: [xp xq] R_b
        const cx<double> y00 = csub(cmul(x00, ecb), cscale(x01, sb));
        const cx<double> y01 = cadd(cmul(x00, esb), cscale(x01, cb));
        const cx<double> y10 = csub(cmul(x10, ecb), cscale(x11, sb));
        const cx<double> y11 = cadd(cmul(x10, esb), cscale(x11, cb));
        // rows: R_a^H
        cx<double> n00 = csub(cmul(cconj(eca), y00), cscale(y10, sa));
        cx<double> n01 = csub(cmul(cconj(eca), y01), cscale(y11, sa));
        cx<double> n10 = cadd(cmul(cconj(esa), y00), cscale(y10, ca));
        cx<double> n11 = cadd(cmul(cconj(esa), y01), cscale(y11, ca));
        if (ka == kb) {
          // diagonal block: Hermitian, the rotated pair's coupling removed
          if (qa >= 0 && sa != 0.0) n10 = z;
          n00.y = 0.0;
          n11.y = 0.0;
          A[ploff(pa) + pa] = n00;
          if (qa >= 0) {
            A[ploff(qa) + pa] = n10;
            A[ploff(qa) + qa] = n11;
          }
        } else {
          pst(A, pa, pb, n00);
          if (qb >= 0) pst(A, pa, qb, n01);
          if (qa >= 0) pst(A, qa, pb, n10);
          if (qa >= 0 && qb >= 0) pst(A, qa, qb, n11);
        }
      }
      // V <- V R: rows p, q of Vt
      for (int it = tid; it < nv; it += NT) {
        const int k = it / n, r = it - k * n;
        const int p = js.p[k], q = js.q[k];
        const double sk = js.s[k];
        if (q >= 0 && sk != 0.0) {
          const double ck = js.c[k];
          const cx<double> ek = js.e[k];
          const cx<double> xp = Vt[p * n + r], xq = Vt[q * n + r];
          Vt[p * n + r] = csub(cmul(xp, cscale(ek, ck)), cscale(xq, sk));
          Vt[q * n + r] = cadd(cmul(xp, cscale(ek, sk)), cscale(xq, ck));
        }
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < n; i += NT) js.ev[i] = A[ploff(i) + i].x;
  for (int i = tid; i < n; i += NT) js.perm[i] = i;
  __syncthreads();
  if (sort) {
    // stable ascending order by rank: index i goes to the number of
    // eigenvalues below it, ties kept in index order (every entry stays an
    // index below n whatever the values)
    for (int i = tid; i < n; i += NT) {
      const double e = js.ev[i];
      int r = 0;
      for (int j = 0; j < n; ++j) {
        const double f = js.ev[j];
        r += (f < e || (f == e && j < i)) ? 1 : 0;
      }
      js.perm[r] = i;
    }
    __syncthreads();
  }
  // Vout[i][j] = component i of eigenvector perm[j]
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e - i * n;
    Vout[e] = Vt[js.perm[j] * n + i];
  }
  double evs = 0.0;
  if (sort && tid < n) evs = js.ev[js.perm[tid]];
  __syncthreads();
  if (sort && tid < n) js.ev[tid] = evs;
  __syncthreads();
  return sweep;
}

constexpr int NDIAG_LDS = 36;   // blocks up to this size keep the five work matrices in LDS
constexpr int NDIAG_MAX = 256;  // largest block hfb_diag_small accepts

// Inverse square root of the Hermitian positive matrix in Y (scaled copy of
// S) by the coupled Newton-Schulz iteration Y <- Y T, Z <- T Z,
// T = (3I - Z Y)/2, which converges to Z = Y0^{-1/2}.  Returns false when it
// has not converged (nearly singular S); the caller then takes the
// eigen-decomposition with the reference's clamp.
template <int NP, int NT>
__device__ __forceinline__ bool newton_schulz(cx<double>*& Y, cx<double>*& Z, cx<double>* P, cx<double>*& Y2, cx<double>*& Z2, int n,
                                              double* red, int& iters) {
  tile_for<NP, NT>(n, [&](int i, int j) { Z[i * n + j] = cx<double>{i == j ? 1.0 : 0.0, 0.0}; });
  __syncthreads();
  int remaining = -1;  // updates left once the residual is below 1e-6
  for (int it = 0; it < 80; ++it) {
    mm<NP, NT, false>(P, Z, Y, n);
    double v = 0.0;
    HFB_TILE_LOOP(NP, NT, n, i, j) {  // residual max|P - I|, then P <- T = (3I - P)/2
      const cx<double> pv = P[i * n + j];
      const double id = i == j ? 1.0 : 0.0;
      v = fmax(v, fabs(pv.x - id) + fabs(pv.y));
      P[i * n + j] = cx<double>{1.5 * id - 0.5 * pv.x, -0.5 * pv.y};
    }
    const double resid = block_max_all<NT>(v, red);
    // quadratic convergence: two updates after 1e-6 reach round-off
    if (remaining < 0 && resid < 1e-6) remaining = 2;
    mm<NP, NT, false>(Y2, Y, P, n);
    mm<NP, NT, false>(Z2, P, Z, n);
    cx<double>* t = Y;
    Y = Y2;
    Y2 = t;
    t = Z;
    Z = Z2;
    Z2 = t;
    iters = it + 1;
    if (remaining > 0 && --remaining == 0) return true;
  }
  return false;
}

// efluct1 = max over the isospins, efluct2 = their mean (static.py:571-576),
// by the second of the two isospin blocks to finish.
__device__ __forceinline__ void diag_finish_arrive(const DiagArgs& a) {
  if (block_arrive(a.sync, gridDim.x) && threadIdx.x == 0) {
    a.efl1[0] = fmax(a.efl1q[0], a.efl1q[1]);
    a.efl2[0] = 0.5 * (a.efl2q[0] + a.efl2q[1]);
  }
}

// Storage of the work matrices B0..B4 and of the Jacobi eigensolver's
// packed matrix Ap and transposed eigenvectors Vt:
//   WM_LDS:  all in LDS (blocks up to NDIAG_LDS states)
//   WM_GL1:  B global, Ap and Vt in LDS (up to 64 states)
//   WM_GL2:  B global, Ap in LDS, Vt global (up to 128 states)
//   WM_GL3:  all global
enum { WM_LDS = 0, WM_GL1 = 1, WM_GL2 = 2, WM_GL3 = 3 };

template <int NP, int NT, int WM>
__global__ void __launch_bounds__(NT) diag_kernel(DiagArgs a) {
  extern __shared__ double lds_raw[];
  // one LDS area for the eigensolver's rotation tables: dense-matrix rounds
  // (work matrices in LDS) or packed-matrix rounds (the larger blocks)
  __shared__ JacobiShared2 js2;
  JacobiShared& js = reinterpret_cast<JacobiShared&>(js2);
  double* const jev = WM == WM_LDS ? js2.ev : js.ev;
  __shared__ double red[32];
  const int iq = blockIdx.x;
  const DiagIso d = a.iso[iq];
  const int n = d.nst;
  if (n <= 0) {
    if (a.construct) diag_finish_arrive(a);
    return;
  }
  const int tid = threadIdx.x;
  const int nm = a.nmax;
  // five work matrices, in LDS for small blocks and in global memory (L2
  // resident) for large ones; each instance sees one kind of pointer only,
  // so the compiler uses LDS or global instructions, never flat ones
  cx<double>* B0;
  if constexpr (WM == WM_LDS)
    B0 = (cx<double>*)lds_raw;
  else
    B0 = d.W;
  cx<double>* const B1 = B0 + n * n;
  cx<double>* const B2 = B1 + n * n;
  cx<double>* const B3 = B2 + n * n;
  cx<double>* const B4 = B3 + n * n;
  cx<double>* Ap;
  cx<double>* Vt;
  if constexpr (WM == WM_LDS || WM == WM_GL3) {
    Ap = B0;
    Vt = B3;
  } else if constexpr (WM == WM_GL1) {
    Ap = (cx<double>*)lds_raw;
    Vt = Ap + ploff(n);
  } else {
    Ap = (cx<double>*)lds_raw;
    Vt = B3;
  }

  cx<double>* Urho = B4;
  cx<double>* U = Urho;
  if (a.mode == DIAG_GIVEN_U) {
    // construct_hfb_matrices with the caller's unitary (static.py:484-577)
    tile_for<NP, NT>(n, [&](int i, int j) { Urho[i * n + j] = d.U[i * n + j]; });
    __syncthreads();
  } else {
    // Loewdin S^{-1/2} (static.py:463-482)
    for (int i = tid; i < n; i += NT) a.sp_norm[d.start + i] = d.S[i * n + i].x;
    load_lower_herm<NP, NT>(B0, d.S, n);
    cx<double>*Y = B0, *Z = B1, *Y2 = B3, *Z2 = B4;
    int ns_iters = 0;
    double cs = 1.0;
    // Near convergence the overlap is S = I + E with a small E: the binomial
    // series (I + E)^{-1/2} = sum_k c_k E^k, c_k = (-1)^k C(2k,k)/4^k, is then
    // exact to round-off in fewer products than the Newton-Schulz iteration
    // (three per step): to degree 4 for |E|_F < 1e-3 (first omitted term below
    // 0.25e-15) in 3 products, to degree 7 for |E|_F < 1e-2 (below 0.2e-16) in
    // 6.  Horner: T = c_{d-1} I + c_d E, then T <- c_k I + E T.  The series
    // works on S itself; Newton-Schulz on S scaled by a Gershgorin bound.
    bool ok;
    {
      double fr = 0.0, dummy = 0.0;
      HFB_TILE_LOOP(NP, NT, n, i, j) {
        const cx<double> v = B0[i * n + j];
        const double e = i == j ? v.x - 1.0 : v.x;
        fr += e * e + v.y * v.y;
      }
      block_sum2_all<NT>(fr, dummy, red);
      const int deg = fr < 1e-6 ? 4 : (fr < 1e-4 ? 7 : 0);
      if (deg > 0) {
        constexpr double ck[8] = {1.0, -0.5, 0.375, -0.3125, 0.2734375, -0.24609375, 0.2255859375, -0.20947265625};
        // E in B0, Horner terms in B2 (products in B3), result Z in B1
        tile_for<NP, NT>(n, [&](int i, int j) {
          if (i == j) B0[i * n + j].x -= 1.0;
          const cx<double> e = B0[i * n + j];
          B2[i * n + j] = cx<double>{(i == j ? ck[deg - 1] : 0.0) + ck[deg] * e.x, ck[deg] * e.y};
        });
        __syncthreads();
        for (int k = deg - 2; k >= 0; --k) {
          mm<NP, NT, false>(B3, B0, B2, n);
          cx<double>* dst = k == 0 ? B1 : B2;
          tile_for<NP, NT>(n, [&](int i, int j) {
            const cx<double> v = B3[i * n + j];
            dst[i * n + j] = cx<double>{v.x + (i == j ? ck[k] : 0.0), v.y};
          });
          __syncthreads();
        }
        Z = B1;
        ok = true;
      } else {
        double rowsum = 0.0;  // Gershgorin bound on the largest eigenvalue
        if (tid < n)
          for (int j = 0; j < n; ++j) rowsum += sqrt(cabs2(B0[tid * n + j]));
        cs = block_max_all<NT>(rowsum, red);
        if (!(cs > 0.0)) cs = 1.0;
        tile_for<NP, NT>(n, [&](int i, int j) { B0[i * n + j] = cscale(B0[i * n + j], 1.0 / cs); });
        __syncthreads();
        ok = newton_schulz<NP, NT>(Y, Z, B2, Y2, Z2, n, red, ns_iters);
      }
    }
    if (ok) {
      const double f = 1.0 / sqrt(cs);
      // Z may live in B4: the scaling is elementwise, so in place is safe
      tile_for<NP, NT>(n, [&](int i, int j) { Urho[i * n + j] = cscale(Z[i * n + j], f); });
      __syncthreads();
    } else {
      if constexpr (WM == WM_LDS) {
        load_lower_herm<NP, NT>(B0, d.S, n);
        jacobi_eigh<NP, NT>(B0, B3, B1, n, js2, false, B2);
      } else {
        load_lower_packed<NT>(Ap, d.S, n);
        jacobi_packed<NT>(Ap, Vt, B1, n, js, false);
      }
      __shared__ double fac[NMAXJ];
      if (tid == 0) {
        double wmax = jev[0];
        for (int i = 1; i < n; ++i) wmax = fmax(wmax, jev[i]);
        const double thr = fmax(1e-12, 2.220446049250313e-16 * wmax * n);
        for (int i = 0; i < n; ++i) fac[i] = 1.0 / sqrt(fmax(jev[i], thr));
      }
      __syncthreads();
      const cx<double>* V = B1;
      tile_for<NP, NT>(n, [&](int i, int j) {  // U = V diag(fac) V^H
        cx<double> acc = {0.0, 0.0};
        for (int k = 0; k < n; ++k) acc = cadd(acc, cscale(cmul(V[i * n + k], cconj(V[j * n + k])), fac[k]));
        Urho[i * n + j] = acc;
      });
      __syncthreads();
    }

    // subspace diagonalisation (static.py:300-303): U <- U_rho U_lam
    if (a.stats && tid == 0) {
      a.stats[2 + iq] = ok ? ns_iters : -1;
      a.stats[iq] = 0;
    }
    if (a.tdiag && d.Lam) {
      int sw;
      if constexpr (WM == WM_LDS) {
        load_lower_herm<NP, NT>(B0, d.Lam, n);
        sw = jacobi_eigh<NP, NT>(B0, B3, B1, n, js2, true, B2);
      } else {
        load_lower_packed<NT>(Ap, d.Lam, n);
        sw = jacobi_packed<NT>(Ap, Vt, B1, n, js, true);
      }
      if (a.stats && tid == 0) a.stats[iq] = sw;
      mm<NP, NT, false>(B3, Urho, B1, n);
      U = B3;
    }
    tile_for<NP, NT>(n, [&](int i, int j) { d.U[i * n + j] = U[i * n + j]; });
    if (a.mode == DIAG_LOEWDIN) return;
  }  // Loewdin and the optional diagonalisation

  if (a.construct) {  // construct_hfb_matrices (static.py:484-577)
    cx<double>* H = B2;
    cx<double>* G = B0;
    tile_for<NP, NT>(n, [&](int i, int j) { B0[i * n + j] = d.Hold[i * n + j]; });
    __syncthreads();
    mm<NP, NT, false>(B1, B0, U, n);
    mm<NP, NT, true>(H, U, B1, n);
    if (a.ipair && d.Gold) {
      tile_for<NP, NT>(n, [&](int i, int j) { B0[i * n + j] = d.Gold[i * n + j]; });
      __syncthreads();
      mm<NP, NT, false>(B1, B0, U, n);
      mm<NP, NT, true>(G, U, B1, n);
    } else {
      tile_for<NP, NT>(n, [&](int i, int j) { G[i * n + j] = {0.0, 0.0}; });
      __syncthreads();
    }
    tile_for<NP, NT>(n, [&](int i, int j) {
      d.Hm[i * nm + j] = H[i * n + j];
      d.Gm[i * nm + j] = G[i * n + j];
    });
    for (int i = tid; i < n; i += NT) {
      a.sp_energy[d.start + i] = H[i * n + i].x;
      if (a.ipair) a.deltaf[d.start + i] = G[i * n + i].x * a.pairwg[d.start + i];
    }
    __syncthreads();
    if (a.tbcs) {
      tile_for<NP, NT>(n, [&](int i, int j) {
        if (i != j) {
          G[i * n + j] = {0.0, 0.0};
          d.Gm[i * nm + j] = {0.0, 0.0};
        }
      });
      __syncthreads();
    }
    // lambda, its parts and the fluctuation measures; lambda+ kept in B1
    double mx = 0.0, sm = 0.0;
    HFB_TILE_LOOP(NP, NT, n, i, j) {
      const int sj = d.start + j, si = d.start + i;
      const double wj = a.wocc[sj] * a.wstates[sj];
      const double uj = a.wguv[sj] * a.pairwg[sj] * a.wstates[sj];
      const double wi = a.wocc[si] * a.wstates[si];
      const double ui = a.wguv[si] * a.pairwg[si] * a.wstates[si];
      const cx<double> lij = csub(cscale(H[i * n + j], wj), cscale(G[i * n + j], uj));
      const cx<double> lji = csub(cscale(H[j * n + i], wi), cscale(G[j * n + i], ui));
      const cx<double> lh = cconj(lji);  // (lambda^H)[i][j]
      const cx<double> asym = cscale(csub(lij, lh), 0.5);
      const cx<double> sym = cscale(cadd(lij, lh), 0.5);
      d.Sym[i * nm + j] = asym;
      d.Ls[i * nm + j] = sym;
      B1[i * n + j] = sym;
      const double aa = cabs2(asym);
      mx = fmax(mx, sqrt(aa));
      sm += aa;
    }
    const double M = block_max_all<NT>(mx, red);
    double Ssum = sm, dummy = 0.0;
    block_sum2_all<NT>(Ssum, dummy, red);
    if (tid == 0) {
      a.efl1q[iq] = M;
      a.efl2q[iq] = sqrt(Ssum / ((double)n * n));
    }
  } else {
    // construct=False: lagrange from the stored lambda (static.py:336-350)
    tile_for<NP, NT>(n, [&](int i, int j) { B1[i * n + j] = d.Ls[i * nm + j]; });
    __syncthreads();
  }
  // M2 = U lambda+ (static.py:557-570); with the caller's unitary the
  // states are already rotated and M2 = lambda+
  if (a.mode == DIAG_GIVEN_U) {
    tile_for<NP, NT>(n, [&](int i, int j) { d.M2[i * n + j] = B1[i * n + j]; });
  } else {
    mm<NP, NT, false>(B0, U, B1, n);
    tile_for<NP, NT>(n, [&](int i, int j) { d.M2[i * n + j] = B0[i * n + j]; });
  }
  if (a.construct) diag_finish_arrive(a);
}

// ---------------------------------------------------------------------------
// Configuration-space sub-iterations of the canonical basis (static.py:920-
// 1070, hfb_natural) and the quasi-particle diagonalisation (static.py:1072-
// 1168, lastdiag), one workgroup, matrices in global memory (L2 resident).
// ---------------------------------------------------------------------------
constexpr int NAT_NT = 256;

struct NatArgs {
  int iq, start, n, nmax, iternat, ipair, phase, e0bas_zero;
  double delstepbas, e0bas, npart, pair_cut, state_cut, softcut;
  cx<double>* hm;  // static arrays, rows of iq, leading dimension nmax
  cx<double>* gm;
  cx<double>* sym;
  cx<double>* ls;
  double *sp_energy, *deltaf, *wocc, *wguv, *pairwg, *wstates;
  double *eferm, *epair, *avdelt, *avdeltv2, *avg;  // (2)
  cx<double>* W;       // 10 n^2 + n complex: cm h g ht gt hcm gcm lam tmp q, tau
  cx<double>* cm_out;  // phase 1: cm
  cx<double>* m2_out;  // phase 1: cm htran
};

__device__ __forceinline__ cx<double> cdiv(cx<double> a, cx<double> b) {
  const double d = b.x * b.x + b.y * b.y;
  return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}

// Q of the QR factorisation of A (n x n, overwritten by the reflectors):
// Householder reflections with LAPACK's zgeqr2/zlarfg conventions (real
// R_kk = beta = -sign(Re alpha) |column|, tau = (beta - alpha)/beta,
// v_k = 1, H_k^H applied from the left) and Q = H_0 H_1 ... H_{n-1}
// accumulated backwards as zung2r does, i.e. the Q of numpy.linalg.qr and
// jnp.linalg.qr.
template <int NT>
__device__ void householder_q(cx<double>* A, cx<double>* Q, cx<double>* tau, int n, double* red) {
  const int tid = threadIdx.x;
  for (int k = 0; k < n; ++k) {
    double xn2 = 0.0, dummy = 0.0;
    for (int i = k + 1 + tid; i < n; i += NT) xn2 += cabs2(A[i * n + k]);
    block_sum2_all<NT>(xn2, dummy, red);
    const cx<double> alpha = A[k * n + k];
    cx<double> tk = {0.0, 0.0};
    if (!(xn2 == 0.0 && alpha.y == 0.0)) {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      const double beta = alpha.x >= 0.0 ? -nrm : nrm;
      tk = {(beta - alpha.x) / beta, -alpha.y / beta};
      const cx<double> sc = cdiv(cx<double>{1.0, 0.0}, cx<double>{alpha.x - beta, alpha.y});
      for (int i = k + 1 + tid; i < n; i += NT) A[i * n + k] = cmul(A[i * n + k], sc);
      __syncthreads();
      if (tid == 0) A[k * n + k] = {beta, 0.0};
    }
    if (tid == 0) tau[k] = tk;
    __syncthreads();
    // columns j > k: A <- (I - conj(tau) v v^H) A
    const cx<double> ct = cconj(tk);
    for (int j = k + 1 + tid; j < n; j += NT) {
      cx<double> w = A[k * n + j];
      for (int i = k + 1; i < n; ++i) w = cadd(w, cmul(cconj(A[i * n + k]), A[i * n + j]));
      const cx<double> f = cmul(ct, w);
      A[k * n + j] = csub(A[k * n + j], f);
      for (int i = k + 1; i < n; ++i) A[i * n + j] = csub(A[i * n + j], cmul(A[i * n + k], f));
    }
    __syncthreads();
  }
  for (int e = tid; e < n * n; e += NT) Q[e] = cx<double>{(e / n) == (e % n) ? 1.0 : 0.0, 0.0};
  __syncthreads();
  for (int k = n - 1; k >= 0; --k) {
    const cx<double> tk = tau[k];
    for (int j = k + tid; j < n; j += NT) {
      cx<double> w = Q[k * n + j];
      for (int i = k + 1; i < n; ++i) w = cadd(w, cmul(cconj(A[i * n + k]), Q[i * n + j]));
      const cx<double> f = cmul(tk, w);
      Q[k * n + j] = csub(Q[k * n + j], f);
      for (int i = k + 1; i < n; ++i) Q[i * n + j] = csub(Q[i * n + j], cmul(A[i * n + k], f));
    }
    __syncthreads();
  }
}

// L (n x n) <- Hermitian part of w_j H_ij - wuv_j G_ij (static.py:1036-1053)
template <int NT>
__device__ void nat_lambda(const NatArgs& a, cx<double>* L, const cx<double>* H, const cx<double>* G) {
  const int n = a.n;
  for (int e = threadIdx.x; e < n * n; e += NT) {
    const int i = e / n, j = e - i * n;
    const int si = a.start + i, sj = a.start + j;
    const double wj = a.wocc[sj] * a.wstates[sj], uj = a.wguv[sj] * a.pairwg[sj] * a.wstates[sj];
    const double wi = a.wocc[si] * a.wstates[si], ui = a.wguv[si] * a.pairwg[si] * a.wstates[si];
    const cx<double> lij = csub(cscale(H[i * n + j], wj), cscale(G[i * n + j], uj));
    const cx<double> lji = csub(cscale(H[j * n + i], wi), cscale(G[j * n + i], ui));
    L[e] = cscale(cadd(lij, cconj(lji)), 0.5);
  }
  __syncthreads();
}

// BCS of the block inside the sub-iterations (static.py:1170-1336): the
// reference's brentq call (rtol=1e-16, below SciPy's minimum) always
// raises, so the Fermi energy is its fallback, the midpoint of block levels
// N-1 and N (or their mean when the block has no such pair).
template <int NT>
__device__ void nat_pair(const NatArgs& a, double* red) {
  const int n = a.n, tid = threadIdx.x;
  const double* e = a.sp_energy + a.start;
  const int nf = (int)llround(a.npart) - 1;
  double ef;
  if (nf >= 0 && nf < n - 1) {
    ef = 0.5 * (e[nf] + e[nf + 1]);
  } else {
    double sm = 0.0, dummy = 0.0;
    for (int i = tid; i < n; i += NT) sm += e[i];
    block_sum2_all<NT>(sm, dummy, red);
    ef = sm / n;
  }
  double v[4] = {0.0, 0.0, 0.0, 0.0};  // sumuv, sumduv, sumv2, sumdv2
  for (int i = tid; i < n; i += NT) {
    const int s = a.start + i;
    const double dl = a.deltaf[s], pw = a.pairwg[s], ws = a.wstates[s];
    const double edif = e[i] - ef, dsq = dl * pw;
    const double eq = fmax(sqrt(edif * edif + dsq * dsq), 1e-20);
    const double v2 = fmin(fmax(0.5 * (1.0 - edif / eq), 1e-6), 1.0 - 1e-6);
    const double uv = sqrt(fmax(v2 * (1.0 - v2), 1e-6));
    double wsn = ws;
    if (a.pair_cut > 0.0)
      a.pairwg[s] = pw / (1.0 + exp((e[i] - (ef + a.pair_cut)) / (a.softcut * a.pair_cut)));
    if (a.state_cut > 0.0) {
      wsn = ws / (1.0 + exp((e[i] - (ef + a.state_cut)) / (a.softcut * a.state_cut)));
      a.wstates[s] = wsn;
    }
    a.wocc[s] = v2;
    a.wguv[s] = uv;
    const double vol = 0.5 * uv * wsn;
    v[0] += vol;
    v[1] += vol * dl;
    v[2] += v2 * wsn;
    v[3] += dl * v2 * wsn;
  }
  block_sum2_all<NT>(v[0], v[1], red);
  block_sum2_all<NT>(v[2], v[3], red);
  if (tid == 0) {
    const int q = a.iq;
    const double suv = fmax(v[0], 1e-20);
    a.eferm[q] = ef;
    a.avdelt[q] = v[1] / suv;
    a.avdeltv2[q] = v[3] / v[2];
    a.epair[q] = v[1];
    a.avg[q] = v[1] / (suv * suv);
  }
  __syncthreads();
}

__global__ void __launch_bounds__(NAT_NT) natural_kernel(NatArgs a) {
  constexpr int NT = NAT_NT;
  __shared__ double red[32];
  __shared__ double damp[NMAXJ];
  const int n = a.n, nm = a.nmax, tid = threadIdx.x;
  const int n2 = n * n;
  cx<double>* const CM = a.W;
  cx<double>* const H = CM + n2;
  cx<double>* const G = H + n2;
  cx<double>* const HT = G + n2;
  cx<double>* const GT = HT + n2;
  cx<double>* const HCM = GT + n2;
  cx<double>* const GCM = HCM + n2;
  cx<double>* const LAM = GCM + n2;
  cx<double>* const TMP = LAM + n2;
  cx<double>* const QQ = TMP + n2;
  cx<double>* const TAU = QQ + n2;
  if (a.phase == 1) {
    // after the full pair of the last sub-iteration: lambda_save, hmatrix,
    // gapmatrix, cm and cm htran for the rotation of psi and lagrange
    nat_lambda<NT>(a, LAM, HT, GT);
    for (int e = tid; e < n2; e += NT) {
      const int i = e / n, j = e - i * n;
      a.ls[i * nm + j] = LAM[e];
      a.hm[i * nm + j] = HT[e];
      a.gm[i * nm + j] = GT[e];
      a.cm_out[e] = CM[e];
    }
    mm<NMAXJ, NT, false>(TMP, CM, HT, n);
    for (int e = tid; e < n2; e += NT) a.m2_out[e] = TMP[e];
    return;
  }
  for (int e = tid; e < n2; e += NT) {
    const int i = e / n, j = e - i * n;
    CM[e] = cx<double>{i == j ? 1.0 : 0.0, 0.0};
    H[e] = a.hm[i * nm + j];
    G[e] = a.gm[i * nm + j];
  }
  // damping per state (static.py:935-944)
  {
    const double* e = a.sp_energy + a.start;
    double mx = -1e300, mn = 1e300;
    for (int i = tid; i < n; i += NT) {
      mx = fmax(mx, e[i]);
      mn = fmin(mn, e[i]);
    }
    mx = block_max_all<NT>(mx, red);
    mn = -block_max_all<NT>(-mn, red);
    for (int i = tid; i < n; i += NT)
      damp[i] = a.e0bas_zero ? a.delstepbas / (mx - mn) : a.delstepbas / (e[i] - e[0] + a.e0bas);
  }
  __syncthreads();
  for (int it = 0; it < a.iternat; ++it) {
    const bool last = it == a.iternat - 1;
    for (int e = tid; e < n2; e += NT) {
      const int j = e / n, k = e - j * n;
      TMP[e] = csub(CM[e], cscale(a.sym[j * nm + k], damp[j]));
    }
    __syncthreads();
    householder_q<NT>(TMP, CM, TAU, n, red);
    mm<NMAXJ, NT, false>(HCM, H, CM, n);
    mm<NMAXJ, NT, false>(GCM, G, CM, n);
    mm<NMAXJ, NT, true>(HT, CM, HCM, n);
    mm<NMAXJ, NT, true>(GT, CM, GCM, n);
    if (a.ipair || last) {
      for (int i = tid; i < n; i += NT) {
        a.sp_energy[a.start + i] = HT[i * n + i].x;
        if (a.ipair) a.deltaf[a.start + i] = GT[i * n + i].x * a.pairwg[a.start + i];
      }
      __syncthreads();
      if (a.ipair && !last) nat_pair<NT>(a, red);
    }
    if (!last) {
      nat_lambda<NT>(a, LAM, HT, GT);
      mm<NMAXJ, NT, false>(TMP, CM, LAM, n);
      for (int e = tid; e < n2; e += NT) {
        const int i = e / n, j = e - i * n;
        const int sj = a.start + j;
        const double wj = a.wocc[sj] * a.wstates[sj], uj = a.wguv[sj] * a.pairwg[sj] * a.wstates[sj];
        a.sym[i * nm + j] = csub(csub(cscale(HCM[e], wj), cscale(GCM[e], uj)), TMP[e]);
      }
      __syncthreads();
    }
  }
}

struct LastArgs {
  int iq, start, n;
  double eferm;
  const cx<double>* hm;  // rows of iq, leading dimension nmax
  const cx<double>* gm;
  int nmax;
  const double* wocc;
  double* qp_energy;
  double* qp_norm;
  cx<double>* W;  // packed n(n+1)/2 + 2 n^2
};

__global__ void __launch_bounds__(NAT_NT) lastdiag_kernel(LastArgs a) {
  constexpr int NT = NAT_NT;
  __shared__ JacobiShared js;
  const int n = a.n, nm = a.nmax, tid = threadIdx.x;
  cx<double>* const Ap = a.W;
  cx<double>* const Vt = Ap + ploff(n);
  cx<double>* const V = Vt + n * n;
  // (h - eferm) xi + Delta eta, its Hermitian part, lower triangle packed
  for (int e = tid; e < ploff(n); e += NT) {
    const int i = prow(e), j = e - ploff(i);
    const double vi = a.wocc[a.start + i], vj = a.wocc[a.start + j];
    const double xi = -sqrt(vi * vj) + sqrt((1.0 - vi) * (1.0 - vj));
    const double eta = sqrt(vi * (1.0 - vj)) + sqrt(vj * (1.0 - vi));
    cx<double> hij = a.hm[i * nm + j], hji = a.hm[j * nm + i];
    if (i == j) {
      hij.x -= a.eferm;
      hji.x -= a.eferm;
    }
    const cx<double> m1 = cadd(cscale(hij, xi), cscale(a.gm[i * nm + j], eta));
    const cx<double> m2 = cadd(cscale(hji, xi), cscale(a.gm[j * nm + i], eta));
    cx<double> v = cscale(cadd(m1, cconj(m2)), 0.5);
    if (i == j) v.y = 0.0;
    Ap[e] = v;
  }
  __syncthreads();
  jacobi_packed<NT>(Ap, Vt, V, n, js, true);
  // lower components: row i of the coefficient matrix has one entry, at the
  // time-reversed partner (global index parity, static.py:1120-1131)
  for (int j = tid; j < n; j += NT) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      const int gi = a.start + i;
      int pj = -1;
      double c = 0.0;
      if (gi % 2 == 0 && i >= 1) {
        pj = i - 1;
        c = -sqrt(a.wocc[gi]);
      } else if (gi % 2 == 1 && i + 1 < n) {
        pj = i + 1;
        c = sqrt(a.wocc[gi]);
      }
      if (pj >= 0) s += c * c * cabs2(V[pj * n + j]);
    }
    a.qp_energy[j] = js.ev[j];
    a.qp_norm[j] = sqrt(s);
  }
}

}  // namespace

// One workgroup per isospin; the work matrices of the larger block decide
// the instance (LDS up to NDIAG_LDS states, global memory above).
static int launch_diag(const DiagArgs& a, int nbig, hipStream_t st) {
#define HFB_DIAG(NPV, NTV, WMV, LDSV)                                                        \
  {                                                                                         \
    if ((LDSV) > 32 * 1024) allow_full_lds((const void*)diag_kernel<NPV, NTV, WMV>);        \
    hipLaunchKernelGGL((diag_kernel<NPV, NTV, WMV>), dim3(2), dim3(NTV), (LDSV), st, a);    \
  }
  const size_t cz = sizeof(cx<double>);
  const size_t tri = (size_t)nbig * (nbig + 1) / 2;
  if (nbig <= NDIAG_LDS) {
    const size_t lds = cz * 5 * (size_t)nbig * nbig;
    if (nbig <= 16)
      HFB_DIAG(16, 256, WM_LDS, lds)
    else if (nbig <= 32)
      HFB_DIAG(32, 512, WM_LDS, lds)
    else
      HFB_DIAG(64, 512, WM_LDS, lds)
  } else if (nbig <= 64) {
    HFB_DIAG(64, 512, WM_GL1, cz * (tri + (size_t)nbig * nbig))
  } else if (nbig <= 128) {
    HFB_DIAG(128, 1024, WM_GL2, cz * tri)
  } else {
    HFB_DIAG(256, 1024, WM_GL3, 0)
  }
#undef HFB_DIAG
  return launch_status();
}

// mats[iq*4 + m]: m=0 S, 1 H_old, 2 lambda_lin (or null), 3 G_old (or null),
// each nst_iq x nst_iq complex; uo[iq*2]: U, uo[iq*2+1]: M2 outputs.
extern "C" int hfb_diag_small(hfb_ctx* c, int nst0, int nst1, int nmax, const double* const* mats, int tdiag, int ipair,
                              int tbcs, int construct, const double* wocc, const double* wstates, const double* wguv,
                              const double* pairwg, double* sp_norm, double* sp_energy, double* deltaf,
                              double* hmatrix, double* gapmatrix, double* symcond, double* lambda_save,
                              double* efluct1q, double* efluct2q, double* efluct1, double* efluct2,
                              double* const* uo, double* work, double* stats, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!mats || !uo || !wocc || !wstates || !wguv || !pairwg || !sp_norm || !sp_energy || !deltaf ||
      !hmatrix || !gapmatrix || !symcond || !lambda_save || !efluct1q || !efluct2q || !efluct1 || !efluct2)
    return HFB_E_ARG;
  if (nst0 < 0 || nst1 < 0 || nmax < nst0 || nmax < nst1) return HFB_E_SIZE;
  if (nst0 > NDIAG_MAX || nst1 > NDIAG_MAX) return HFB_E_SIZE;
  const int nbig0 = nst0 > nst1 ? nst0 : nst1;
  if (nbig0 > NDIAG_LDS && !work) return HFB_E_ARG;
  DiagArgs a{};
  a.nmax = nmax;
  const int nsts[2] = {nst0, nst1};
  const int64_t nm2 = (int64_t)nmax * nmax;
  for (int iq = 0; iq < 2; ++iq) {
    DiagIso& d = a.iso[iq];
    d.nst = nsts[iq];
    d.start = iq == 0 ? 0 : nst0;
    d.S = (const cx<double>*)mats[iq * 4 + 0];
    d.Hold = (const cx<double>*)mats[iq * 4 + 1];
    d.Lam = (const cx<double>*)mats[iq * 4 + 2];
    d.Gold = (const cx<double>*)mats[iq * 4 + 3];
    if (d.nst > 0 && (!d.S || !d.Hold || (tdiag && !d.Lam) || (ipair && construct && !d.Gold))) return HFB_E_ARG;
    d.U = (cx<double>*)uo[iq * 2 + 0];
    d.M2 = (cx<double>*)uo[iq * 2 + 1];
    if (d.nst > 0 && (!d.U || !d.M2)) return HFB_E_ARG;
    d.W = work ? (cx<double>*)work + (int64_t)iq * 5 * nm2 : nullptr;
    d.Hm = (cx<double>*)hmatrix + iq * nm2;
    d.Gm = (cx<double>*)gapmatrix + iq * nm2;
    d.Sym = (cx<double>*)symcond + iq * nm2;
    d.Ls = (cx<double>*)lambda_save + iq * nm2;
  }
  a.wocc = wocc;
  a.wstates = wstates;
  a.wguv = wguv;
  a.pairwg = pairwg;
  a.sp_norm = sp_norm;
  a.sp_energy = sp_energy;
  a.deltaf = deltaf;
  a.efl1q = efluct1q;
  a.efl2q = efluct2q;
  a.efl1 = efluct1;
  a.efl2 = efluct2;
  a.sync = c->sync + HFB_SYNC_DIAG;
  a.tdiag = tdiag;
  a.ipair = ipair;
  a.tbcs = tbcs;
  a.construct = construct;
  a.stats = stats;
  return launch_diag(a, nst0 > nst1 ? nst0 : nst1, (hipStream_t)stream);
}

// Loewdin S^{-1/2} of one block (static.py:463-482): U (nst x nst complex)
// and sp_norm = Re diag S; work: null up to NDIAG_LDS states, else 5 nst^2
// complex doubles.
extern "C" int hfb_loewdin(int nst, const double* S, double* U, double* sp_norm, double* work, void* stream) {
  if (!S || !U || !sp_norm) return HFB_E_ARG;
  if (nst <= 0 || nst > NDIAG_MAX) return HFB_E_SIZE;
  if (nst > NDIAG_LDS && !work) return HFB_E_ARG;
  DiagArgs a{};
  a.nmax = nst;
  DiagIso& d = a.iso[0];
  d.nst = nst;
  d.S = (const cx<double>*)S;
  d.U = (cx<double>*)U;
  d.W = (cx<double>*)work;
  a.sp_norm = sp_norm;
  a.mode = DIAG_LOEWDIN;
  return launch_diag(a, nst, (hipStream_t)stream);
}

// construct_hfb_matrices (static.py:484-577) of isospin iq with the
// caller's unitary U (states start .. start+nst-1 already rotated):
// H = U^H H_old U, G likewise (ipair), hmatrix/gapmatrix/symcond/lambda_save
// rows of iq, sp_energy, deltaf, efluct1q/2q[iq], efluct1/2 over both
// isospins, and lam = lambda+ (nst x nst) with lagrange = psi lam.
extern "C" int hfb_hfb_matrices(hfb_ctx* c, int iq, int start, int nst, int nmax, const double* U, const double* Hold,
                                const double* Gold, int ipair, int tbcs, const double* wocc, const double* wstates,
                                const double* wguv, const double* pairwg, double* sp_energy, double* deltaf,
                                double* hmatrix, double* gapmatrix, double* symcond, double* lambda_save,
                                double* efluct1q, double* efluct2q, double* efluct1, double* efluct2, double* lam,
                                double* work, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!U || !Hold || (ipair && !Gold) || !wocc || !wstates || !wguv || !pairwg || !sp_energy || !deltaf ||
      !hmatrix || !gapmatrix || !symcond || !lambda_save || !efluct1q || !efluct2q || !efluct1 || !efluct2 || !lam)
    return HFB_E_ARG;
  if (iq < 0 || iq > 1 || start < 0 || nst <= 0 || nst > nmax || nst > NDIAG_MAX) return HFB_E_SIZE;
  if (nst > NDIAG_LDS && !work) return HFB_E_ARG;
  DiagArgs a{};
  a.nmax = nmax;
  DiagIso& d = a.iso[iq];
  d.nst = nst;
  d.start = start;
  d.U = (cx<double>*)U;
  d.Hold = (const cx<double>*)Hold;
  d.Gold = ipair ? (const cx<double>*)Gold : nullptr;
  d.M2 = (cx<double>*)lam;
  d.W = (cx<double>*)work;
  const int64_t nm2 = (int64_t)nmax * nmax;
  d.Hm = (cx<double>*)hmatrix + iq * nm2;
  d.Gm = (cx<double>*)gapmatrix + iq * nm2;
  d.Sym = (cx<double>*)symcond + iq * nm2;
  d.Ls = (cx<double>*)lambda_save + iq * nm2;
  a.wocc = wocc;
  a.wstates = wstates;
  a.wguv = wguv;
  a.pairwg = pairwg;
  a.sp_energy = sp_energy;
  a.deltaf = deltaf;
  a.efl1q = efluct1q;
  a.efl2q = efluct2q;
  a.efl1 = efluct1;
  a.efl2 = efluct2;
  a.sync = c->sync + HFB_SYNC_DIAG;
  a.ipair = ipair;
  a.tbcs = tbcs;
  a.construct = 1;
  a.mode = DIAG_GIVEN_U;
  return launch_diag(a, nst, (hipStream_t)stream);
}

// Sub-iterations of hfb_natural (static.py:920-1070) for isospin iq, states
// start .. start+nst-1.  phase 0: the iternat sub-iterations up to the
// full pair of the last one (sp_energy, deltaf, wocc, ..., symcond and the
// pairing results of iq updated on the way); the caller then runs the full
// pair (hfb_bcs) when ipair; phase 1: lambda_save, hmatrix, gapmatrix of iq,
// cm_out = cm and m2_out = cm htran (psi <- psi cm, lagrange <- psi cm htran
// by hfb_rotate).  work: 10 nst^2 + nst complex doubles, kept between the
// phases.
extern "C" int hfb_natural(int phase, int iq, int start, int nst, int nmax, int iternat, int ipair, double delstepbas,
                           double e0bas, double npart, double pair_cut, double state_cut, double softcut,
                           double* hmatrix, double* gapmatrix, double* symcond, double* lambda_save,
                           double* sp_energy, double* deltaf, double* wocc, double* wguv, double* pairwg,
                           double* wstates, double* eferm, double* epair, double* avdelt, double* avdeltv2,
                           double* avg, double* work, double* cm_out, double* m2_out, void* stream) {
  if (phase != 0 && phase != 1) return HFB_E_ARG;
  if (iq < 0 || iq > 1 || start < 0 || nst <= 0 || nst > nmax || nst > NDIAG_MAX || iternat <= 0) return HFB_E_SIZE;
  if (!hmatrix || !gapmatrix || !symcond || !lambda_save || !sp_energy || !deltaf || !wocc || !wguv || !pairwg ||
      !wstates || !eferm || !epair || !avdelt || !avdeltv2 || !avg || !work)
    return HFB_E_ARG;
  if (phase == 1 && (!cm_out || !m2_out)) return HFB_E_ARG;
  NatArgs a{};
  a.iq = iq;
  a.start = start;
  a.n = nst;
  a.nmax = nmax;
  a.iternat = iternat;
  a.ipair = ipair;
  a.phase = phase;
  a.e0bas_zero = e0bas == 0.0;
  a.delstepbas = delstepbas;
  a.e0bas = e0bas;
  a.npart = npart;
  a.pair_cut = pair_cut;
  a.state_cut = state_cut;
  a.softcut = softcut;
  const int64_t nm2 = (int64_t)nmax * nmax;
  a.hm = (cx<double>*)hmatrix + iq * nm2;
  a.gm = (cx<double>*)gapmatrix + iq * nm2;
  a.sym = (cx<double>*)symcond + iq * nm2;
  a.ls = (cx<double>*)lambda_save + iq * nm2;
  a.sp_energy = sp_energy;
  a.deltaf = deltaf;
  a.wocc = wocc;
  a.wguv = wguv;
  a.pairwg = pairwg;
  a.wstates = wstates;
  a.eferm = eferm;
  a.epair = epair;
  a.avdelt = avdelt;
  a.avdeltv2 = avdeltv2;
  a.avg = avg;
  a.W = (cx<double>*)work;
  a.cm_out = (cx<double>*)cm_out;
  a.m2_out = (cx<double>*)m2_out;
  hipLaunchKernelGGL(natural_kernel, dim3(1), dim3(NAT_NT), 0, (hipStream_t)stream, a);
  return launch_status();
}

// lastdiag (static.py:1072-1168) of isospin iq over states start ..
// start+nst-1: quasi-particle energies (ascending) and the norms of their
// lower components.  work: nst (nst + 1) / 2 + 2 nst^2 complex doubles.
extern "C" int hfb_lastdiag(int iq, int start, int nst, int nmax, double eferm, const double* hmatrix,
                            const double* gapmatrix, const double* wocc, double* qp_energy, double* qp_norm,
                            double* work, void* stream) {
  if (iq < 0 || iq > 1 || start < 0 || nst <= 0 || nst > nmax || nst > NDIAG_MAX) return HFB_E_SIZE;
  if (!hmatrix || !gapmatrix || !wocc || !qp_energy || !qp_norm || !work) return HFB_E_ARG;
  LastArgs a{};
  a.iq = iq;
  a.start = start;
  a.n = nst;
  a.nmax = nmax;
  a.eferm = eferm;
  const int64_t nm2 = (int64_t)nmax * nmax;
  a.hm = (const cx<double>*)hmatrix + iq * nm2;
  a.gm = (const cx<double>*)gapmatrix + iq * nm2;
  a.wocc = wocc;
  a.qp_energy = qp_energy;
  a.qp_norm = qp_norm;
  a.W = (cx<double>*)work;
  hipLaunchKernelGGL(lastdiag_kernel, dim3(1), dim3(NAT_NT), 0, (hipStream_t)stream, a);
  return launch_status();
}
