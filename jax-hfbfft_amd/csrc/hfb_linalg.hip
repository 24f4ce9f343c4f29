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

#include <algorithm>
#include <type_traits>

#include "hfb_ctx.h"

using namespace hfb;

namespace {

constexpr int NMAXJ = 256;
constexpr int NDIAG_LDS = 36;   // blocks up to this size keep the five work matrices in LDS
constexpr int NDIAG_MAX = 256;  // largest block hfb_diag_small accepts

// Jacobi stops when the squared off-diagonal Frobenius norm is below
// tol2 times the squared diagonal one.  The default JACOBI_TOL2 (relative
// off-diagonal norm 1e-13) is the accuracy of FP64 subspace matrices (sums
// of K = 2N products, about sqrt(K) eps = 6e-14 relative at 32^3), so
// further sweeps only polish round-off; measured at Ca40, two sweeps reach
// 1e-28 .. 1e-26 in the steady iterations, the third went on to 1e-36.
// Callers whose matrices come from FP32 states (entries accurate to about
// 1e-7) pass a looser bound for the lambda_lin eigensolve.
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
  // a subspace-diagonalising call with its blocks in LDS on two workgroups
  // per isospin: blockIdx.x = iq + 2 role; role 1 computes U_rho (Loewdin)
  // and hands it over through U, role 0 runs the eigensolver of
  // lambda_lin (independent of U_rho, static.py:295-302), then takes U_rho,
  // forms U = U_rho U_lam and constructs the HFB matrices
  int split;
  double tol2;  // the lambda_lin eigensolve stops at off-diagonal^2 <= tol2 diagonal^2 (Frobenius)
  int mode;  // DIAG_FULL, DIAG_LOEWDIN (S -> U only) or DIAG_GIVEN_U (construct from U in d.U)
  // the wide path (blocks above NDIAG_LDS states, DIAG_FULL without the
  // diagonalisation): per isospin 4 or 7 = the binomial series of that
  // degree runs on the wide kernels, 1 = empty block, 0 = diag_kernel takes
  // the block (Newton-Schulz or eigh); null: diag_kernel takes every block
  unsigned int* ctl;
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

// C = op(A) B for n x n complex matrices (leading dimension n; C distinct
// from A and B); op(A) = A or A^H; LDS or global (L2-resident) operands.
// On the FP64 matrix cores: each wavefront takes 16 x 16 output tiles in
// turn and, per k-step of 4, issues four real products into the real and
// imaginary accumulators, Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br
// (v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[l / 16 + 4 r][l % 16], r = 0..3).  Entries
// beyond n enter as zeros; the next k-step's operands load during this
// step's products.  Each operand entry is one 16-byte load per lane and
// k-step, so the products are not bound by LDS bandwidth.
__device__ __forceinline__ void mac(cx<double>& c, const cx<double> a, const cx<double> b) {
  c.x = fma(a.x, b.x, fma(-a.y, b.y, c.x));
  c.y = fma(a.x, b.y, fma(a.y, b.x, c.y));
}
typedef double mm_d4 __attribute__((ext_vector_type(4)));
template <bool CONJ_A>
__device__ __forceinline__ void mm_operands(const cx<double>* A, const cx<double>* B, int n, int i, int j, int k,
                                            cx<double>& a, cx<double>& b) {
  const bool ka = i < n && k < n, kb = k < n && j < n;
  const int ic = ka ? i : 0, kc = k < n ? k : 0, jc = kb ? j : 0;
  const cx<double> av = CONJ_A ? cconj(A[kc * n + ic]) : A[ic * n + kc];
  const cx<double> bv = B[kc * n + jc];
  a = ka ? av : cx<double>{0.0, 0.0};
  b = kb ? bv : cx<double>{0.0, 0.0};
}

template <int NP, int NT, bool CONJ_A>
__device__ __forceinline__ void mm(cx<double>* __restrict__ C, const cx<double>* A, const cx<double>* B, int n) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nt = (n + 15) >> 4;
  for (int tile = wave; tile < nt * nt; tile += NT / 64) {
    const int ti = tile / nt, tj = tile - ti * nt;
    const int i = ti * 16 + li, j = tj * 16 + li;
    mm_d4 cr = {0.0, 0.0, 0.0, 0.0}, ci = {0.0, 0.0, 0.0, 0.0};
    cx<double> a, b;
    mm_operands<CONJ_A>(A, B, n, i, j, lk, a, b);
    for (int k0 = 0; k0 < n; k0 += 4) {
      cx<double> an, bn;
      mm_operands<CONJ_A>(A, B, n, i, j, k0 + 4 + lk, an, bn);
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.x, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.y, ci, 0, 0, 0);
      cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.y, b.y, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.x, ci, 0, 0, 0);
      a = an;
      b = bn;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = ti * 16 + lk + 4 * r;
      if (row < n && j < n) C[row * n + j] = cx<double>{cr[r], ci[r]};
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

// Eigenvalues ev (n) into ascending order with the columns of V (n x n,
// leading dimension n) following them; T is n x n scratch.  Stable order
// by rank: index i goes to the number of eigenvalues below it, ties kept in
// index order (perm starts as the identity, so that every entry stays an
// index below n whatever the values).
template <int NP, int NT>
__device__ __forceinline__ void sort_eigen(cx<double>* V, cx<double>* T, int n, double* ev, int* perm) {
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += NT) perm[i] = i;
  __syncthreads();
  for (int i = tid; i < n; i += NT) {
    const double e = ev[i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double f = ev[j];
      r += (f < e || (f == e && j < i)) ? 1 : 0;
    }
    perm[r] = i;
  }
  __syncthreads();
  tile_for<NP, NT>(n, [&](int i, int j) { T[i * n + j] = V[i * n + perm[j]]; });
  __syncthreads();
  tile_for<NP, NT>(n, [&](int i, int j) { V[i * n + j] = T[i * n + j]; });
  double evs = 0.0;
  if (tid < n) evs = ev[perm[tid]];
  __syncthreads();
  if (tid < n) ev[tid] = evs;
  __syncthreads();
}

// LDS state of the wavefront Jacobi (blocks up to NDIAG_LDS states).
struct JacobiSmall {
  double ev[NDIAG_LDS];
  int perm[NDIAG_LDS];
  double part[2][3][2];  // per sweep parity and rounds wavefront (at most 3): off, dia
};

// Pair k of round t of the round-robin ordering over m (even) indices:
// index 0 stays, the others move by one place per round; p < q.
__device__ __forceinline__ void rr_pair(int k, int t, int m, int& p, int& q) {
  const int m1 = m - 1;
  int xa = k - 1 + t;
  xa = xa >= m1 ? xa - m1 : xa;
  int xb = m1 - k - 1 + t;
  xb = xb >= m1 ? xb - m1 : xb;
  const int a0 = k == 0 ? 0 : xa + 1, b0 = xb + 1;
  p = a0 < b0 ? a0 : b0;
  q = a0 < b0 ? b0 : a0;
}

// Rotation R = [[c, sg], [-conj(sg), c]] (c real, c^2 + |sg|^2 = 1) that
// zeroes entry (p, q) of R^H [[app, apq], [conj(apq), aqq]] R, the smaller
// of the two angles: with d = (aqq - app)/2, rho = sqrt(d^2 + |apq|^2) and
// w = rho + |d|, tan(theta) = sign(d) |apq| / w, so c = w / sqrt(2 rho w)
// and sg = sign(d) apq / sqrt(2 rho w) (since w^2 + |apq|^2 = 2 rho w).
// Two inverse square roots and no division.  Entries negligible against
// the diagonal (below 1e-18 relative) are not rotated.  Branch-free (the
// unrotated case runs the same arithmetic on a safe operand and is selected
// away), so that the loads of a round issue together and the rotations of
// several pairs interleave.
__device__ __forceinline__ void jw_rot(double app, double aqq, cx<double> apq, double& c, cx<double>& sg) {
  const double r2 = cabs2(apq), den = fabs(app) + fabs(aqq);
  const bool rot = r2 > 0.0 && r2 > 1e-36 * den * den;
  const double d = 0.5 * (aqq - app);
  const double rho2 = rot ? fma(d, d, r2) : 1.0;
  const double rho = rho2 * rsq_full(rho2);
  const double w = rho + fabs(d);
  const double f = rsq_full(2.0 * rho * w);
  const double g = d >= 0.0 ? f : -f;
  c = rot ? w * f : 1.0;
  sg = rot ? cx<double>{apq.x * g, apq.y * g} : cx<double>{0.0, 0.0};
}

__device__ __forceinline__ void wave_sync() {
  // orders this wavefront's LDS accesses for the compiler; the hardware
  // executes one wavefront's LDS instructions in order
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS of the wavefront Jacobi after the five n x n work matrices: the
// rotation log, two sweeps (double buffer) of (m - 1) rounds x m/2 pairs,
// cosines then the complex sines; for odd n also the second matrix buffer
// (m x m; for even n it is B0).
__host__ __device__ __forceinline__ int jw_m(int n) { return n + (n & 1); }
__host__ __device__ __forceinline__ size_t jw_log_bytes(int n) {
  const int m = jw_m(n);
  return (size_t)2 * (m - 1) * (m / 2) * (sizeof(double) + 2 * sizeof(double));
}
__host__ __device__ __forceinline__ size_t jw_extra_bytes(int n) {
  const int m = jw_m(n);
  return jw_log_bytes(n) + ((n & 1) ? (size_t)m * m * 2 * sizeof(double) : 0);
}
// wavefronts that run the rounds: one 2x2 block per lane
__host__ __device__ __forceinline__ int jw_nrw(int n) {
  const int np = jw_m(n) / 2;
  return (np * (np + 1) / 2 + 63) / 64;
}

// The 2x2 block (pair a, pair b) = (ka, kb) of one Jacobi round: R_a^H X
// R_b of the matrix A (m x m), written with its mirror into An (a round
// rewrites every entry once, so with A and An distinct the blocks of one
// round never write what another still reads); the lane of a diagonal
// block logs its rotation (lcr, lsr: this round's slots).  ka = -1: no
// block (the lane works on block (0, 0) and writes nothing).  With `last`,
// off/dia receive the block's norms (entries on and off the diagonal summed
// apart, no cancellation).  Loads first, no branches before the writes.
__device__ __forceinline__ void jw_block(const cx<double>* __restrict__ A, cx<double>* __restrict__ An,
                                         double* __restrict__ lcr, cx<double>* __restrict__ lsr, int m, int t, int ka,
                                         int kb, bool last, double& off, double& dia) {
  const bool own = ka >= 0;
  const int ia = own ? ka : 0, ib = own ? kb : 0;
  const bool dg = own && ia == ib;
  int pa, qa, pb, qb;
  rr_pair(ia, t, m, pa, qa);
  rr_pair(ib, t, m, pb, qb);
  const cx<double> x00 = A[pa * m + pb], x01 = A[pa * m + qb];
  const cx<double> x10 = A[qa * m + pb], x11 = A[qa * m + qb];
  const double app = A[pa * m + pa].x, aqq = A[qa * m + qa].x;
  const cx<double> apq = A[pa * m + qa];
  const double bpp = A[pb * m + pb].x, bqq = A[qb * m + qb].x;
  const cx<double> bpq = A[pb * m + qb];
  double ca, cb;
  cx<double> sa, sb;
  jw_rot(app, aqq, apq, ca, sa);
  jw_rot(bpp, bqq, bpq, cb, sb);
  // columns: [Y_p Y_q] = [X_p X_q] R_b
  const cx<double> y00 = csub(cscale(x00, cb), cmul(x01, cconj(sb)));
  const cx<double> y01 = cadd(cmul(x00, sb), cscale(x01, cb));
  const cx<double> y10 = csub(cscale(x10, cb), cmul(x11, cconj(sb)));
  const cx<double> y11 = cadd(cmul(x10, sb), cscale(x11, cb));
  // rows: R_a^H = [[c, -sg], [conj(sg), c]]
  cx<double> n00 = csub(cscale(y00, ca), cmul(sa, y10));
  cx<double> n01 = csub(cscale(y01, ca), cmul(sa, y11));
  cx<double> n10 = cadd(cmul(cconj(sa), y00), cscale(y10, ca));
  cx<double> n11 = cadd(cmul(cconj(sa), y01), cscale(y11, ca));
  // the rotated pair's diagonal block is exactly diagonal
  const bool zap = dg && (sa.x != 0.0 || sa.y != 0.0);
  n01 = zap ? cx<double>{0.0, 0.0} : n01;
  n10 = zap ? cx<double>{0.0, 0.0} : n10;
  n00.y = zap ? 0.0 : n00.y;
  n11.y = zap ? 0.0 : n11.y;
  if (last) {
    const double d0 = cabs2(n00) + cabs2(n11), o0 = cabs2(n01) + cabs2(n10);
    off = own ? (dg ? o0 : 2.0 * (d0 + o0)) : 0.0;
    dia = dg ? d0 : 0.0;
  }
  if (own) {
    An[pa * m + pb] = n00;
    An[pa * m + qb] = n01;
    An[qa * m + pb] = n10;
    An[qa * m + qb] = n11;
  }
  if (own && !dg) {
    An[pb * m + pa] = cconj(n00);
    An[qb * m + pa] = cconj(n01);
    An[pb * m + qa] = cconj(n10);
    An[qb * m + qa] = cconj(n11);
  }
  if (dg) {
    lcr[ia] = ca;
    lsr[ia] = sa;
  }
}

// One sweep (rounds R0 .. R0 + m - 2) of the wavefront Jacobi when one
// wavefront runs the rounds: round R reads buffer R & 1 and writes the
// other; the wavefront's LDS accesses execute in order, so the next round
// sees this round's writes without a barrier.  Rotations go to the sweep's
// log (lcb, lsb: (m - 1) rounds x m/2 pairs); off/dia receive this lane's
// share of the norms of the sweep's last matrix.
__device__ __forceinline__ void jw_sweep(cx<double>* __restrict__ Ab0, cx<double>* __restrict__ Ab1,
                                         double* __restrict__ lcb, cx<double>* __restrict__ lsb, int m, int R0,
                                         int ka, int kb, double& off, double& dia) {
  const int nrnd = m - 1, npair = m >> 1;
  for (int t = 0; t < nrnd; ++t) {
    const int R = R0 + t;
    jw_block((R & 1) ? Ab1 : Ab0, (R & 1) ? Ab0 : Ab1, lcb + t * npair, lsb + t * npair, m, t, ka, kb,
             t == nrnd - 1, off, dia);
    wave_sync();
  }
}

// V <- V R for one round t (rotations in lcr, lsr); row r of V belongs to
// eigenvector wavefront w = r % nvw, whose lanes take its (row, pair) items.
__device__ __forceinline__ void jw_vround(cx<double>* __restrict__ V, const double* __restrict__ lcr,
                                          const cx<double>* __restrict__ lsr, int n, int m, int t, int w, int nvw,
                                          int lane) {
  const int npair = m >> 1;
  const int nrow = w < n ? (n - w + nvw - 1) / nvw : 0;
  const int nitem = nrow * npair;
  for (int it = lane; it < nitem; it += 64) {
    const int lr = it / npair, k = it - lr * npair;
    const int r = w + nvw * lr;
    int p, q;
    rr_pair(k, t, m, p, q);
    if (q < n) {
      const double c = lcr[k];
      const cx<double> sg = lsr[k];
      const cx<double> vp = V[r * n + p], vq = V[r * n + q];
      V[r * n + p] = csub(cscale(vp, c), cmul(vq, cconj(sg)));
      V[r * n + q] = cadd(cmul(vp, sg), cscale(vq, c));
    }
  }
}

// Eigen-decomposition of the Hermitian matrix whose lower triangle is M
// (global, n x n, leading dimension n) for blocks of up to NDIAG_LDS states
// with their work matrices in LDS: V (n x n) gets the eigenvectors as
// columns, js.ev the eigenvalues, in ascending order when `sort` (T, n x n,
// is scratch).  A0 and A1 (LDS, m x m with m = n rounded up to even; the
// extra index is a zero row and column, whose rotations are the identity)
// hold the whole Hermitian matrix, alternately.  Parallel cyclic Jacobi
// with the pair ordering of rr_pair, A <- R^H A R per round with R
// block-diagonal over the m/2 pairs.  The rounds run in the first
// jw_nrw(n) wavefronts without workgroup barriers (jw_sweep): their lanes
// own the 2x2 blocks (pair a >= pair b) of a round, and each lane reads
// its block and the diagonal blocks of its two pairs, computes both
// rotations itself (jw_rot) and writes R_a^H X R_b and its mirror.  The
// other wavefronts apply the logged rotations of the previous sweep to V
// (a row of V stays in one wavefront, whose LDS accesses execute in order)
// while the first ones run the next sweep; a barrier separates sweeps.  The
// off-diagonal norm of a sweep's last round decides whether another sweep
// follows.
template <int NP, int NT>
__device__ __forceinline__ int jacobi_wave(const cx<double>* __restrict__ M, cx<double>* __restrict__ A0,
                                           cx<double>* __restrict__ A1, cx<double>* __restrict__ V,
                                           double* __restrict__ lc, cx<double>* __restrict__ ls, int n,
                                           JacobiSmall& js, bool sort, cx<double>* T, double tol2 = JACOBI_TOL2) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (n == 1) {
    if (tid == 0) {
      js.ev[0] = M[0].x;
      V[0] = {1.0, 0.0};
    }
    __syncthreads();
    return 0;
  }
  const int m = jw_m(n), npair = m >> 1, nrnd = m - 1, nblk = npair * (npair + 1) / 2;
  const int na = jw_nrw(n);  // rounds wavefronts (the host checks na < NT / 64)
  const int lsz = nrnd * npair;  // log entries per sweep
  for (int e = tid; e < m * m; e += NT) {
    const int i = e / m, j = e - i * m;
    cx<double> v = {0.0, 0.0};
    if (i < n && j < n) {
      if (i > j)
        v = M[i * n + j];
      else if (i < j)
        v = cconj(M[j * n + i]);
      else
        v = {M[i * n + i].x, 0.0};
    }
    A0[e] = v;
  }
  tile_for<NP, NT>(n, [&](int i, int j) { V[i * n + j] = cx<double>{i == j ? 1.0 : 0.0, 0.0}; });
  __syncthreads();

  // this lane's block (ka >= kb) of the lower block triangle
  int ka = -1, kb = 0;
  if (wave < na) {
    const int blk = lane + 64 * wave;
    int a = (int)((sqrtf(8.0f * blk + 1.0f) - 1.0f) * 0.5f);
    while (a * (a + 1) / 2 > blk) --a;
    while ((a + 1) * (a + 2) / 2 <= blk) ++a;
    if (blk < nblk) {
      ka = a;
      kb = blk - a * (a + 1) / 2;
    }
    // norms of the initial matrix
    double off = 0.0, dia = 0.0;
    if (ka >= 0) {
      int pa, qa, pb, qb;
      rr_pair(ka, 0, m, pa, qa);
      rr_pair(kb, 0, m, pb, qb);
      const double d0 = cabs2(A0[pa * m + pb]) + cabs2(A0[qa * m + qb]);
      const double o0 = cabs2(A0[pa * m + qb]) + cabs2(A0[qa * m + pb]);
      off = ka != kb ? 2.0 * (d0 + o0) : o0;
      dia = ka != kb ? 0.0 : d0;
    }
#pragma unroll
    for (int h = 1; h < 64; h <<= 1) {
      off += __shfl_xor(off, h, 64);
      dia += __shfl_xor(dia, h, 64);
    }
    if (lane == 0) {
      js.part[0][wave][0] = off;
      js.part[0][wave][1] = dia;
    }
  }
  __syncthreads();
  const int nvw = NT / 64 - na;  // eigenvector wavefronts
  const int w = wave - na;
  int sweep = 0;
  for (;; ++sweep) {
    const int buf = sweep & 1;
    // the norms of the current matrix (the previous sweep's last round):
    // every wavefront takes the same decision
    double off = 0.0, dia = 0.0;
    for (int k = 0; k < na; ++k) {
      off += js.part[buf][k][0];
      dia += js.part[buf][k][1];
    }
    const bool stop = sweep >= 40 || off <= tol2 * dia || off == 0.0;
    double o = 0.0, d = 0.0;
    if (na == 1) {
      // one rounds wavefront, no barriers inside a sweep: the eigenvector
      // wavefronts apply the previous sweep's logged rotations meanwhile
      if (wave == 0) {
        if (!stop) jw_sweep(A0, A1, lc + buf * lsz, ls + buf * lsz, m, sweep * nrnd, ka, kb, o, d);
      } else if (sweep > 0) {
        for (int t = 0; t < nrnd; ++t) {
          jw_vround(V, lc + (buf ^ 1) * lsz + t * npair, ls + (buf ^ 1) * lsz + t * npair, n, m, t, w, nvw, lane);
          wave_sync();
        }
      }
    } else if (!stop) {
      // several rounds wavefronts: a workgroup barrier after every round;
      // the eigenvector wavefronts apply the previous round meanwhile (the
      // first round of a sweep: the previous sweep's last round), from the
      // two-slot log of the first 2 m/2 entries
      for (int t = 0; t < nrnd; ++t) {
        const int R = sweep * nrnd + t;
        if (wave < na) {
          jw_block((R & 1) ? A1 : A0, (R & 1) ? A0 : A1, lc + (R & 1) * npair, ls + (R & 1) * npair, m, t, ka, kb,
                   t == nrnd - 1, o, d);
        } else if (R > 0) {
          const int Rp = R - 1;
          jw_vround(V, lc + (Rp & 1) * npair, ls + (Rp & 1) * npair, n, m, Rp % nrnd, w, nvw, lane);
        }
        __syncthreads();
      }
    } else if (sweep > 0 && wave >= na) {
      // converged: the last round's rotations
      const int Rp = sweep * nrnd - 1;
      jw_vround(V, lc + (Rp & 1) * npair, ls + (Rp & 1) * npair, n, m, Rp % nrnd, w, nvw, lane);
    }
    if (wave < na && !stop) {
#pragma unroll
      for (int h = 1; h < 64; h <<= 1) {
        o += __shfl_xor(o, h, 64);
        d += __shfl_xor(d, h, 64);
      }
      if (lane == 0) {
        js.part[buf ^ 1][wave][0] = o;
        js.part[buf ^ 1][wave][1] = d;
      }
    }
    __syncthreads();
    if (stop) break;
  }
  // the final matrix: buffer (sweeps * rounds) & 1
  const cx<double>* const Af = ((sweep * nrnd) & 1) ? A1 : A0;
  for (int i = tid; i < n; i += NT) js.ev[i] = Af[i * m + i].x;
  __syncthreads();
  if (sort) sort_eigen<NP, NT>(V, T, n, js.ev, js.perm);
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
                                             cx<double>* __restrict__ Vout, int n, JacobiShared& js, bool sort,
                                             double tol2 = JACOBI_TOL2) {
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
    if (off <= tol2 * dia || off == 0.0) break;
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
  if (block_arrive(a.sync, 2u) && threadIdx.x == 0) {
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
  // the eigensolver's LDS state: the wavefront Jacobi (blocks in LDS) or
  // the packed-triangle rounds (the larger blocks)
  using JS = std::conditional_t<WM == WM_LDS, JacobiSmall, JacobiShared>;
  __shared__ JS jsh;
  double* const jev = jsh.ev;
  __shared__ double red[32];
  const int iq = blockIdx.x & 1;
  const int role = blockIdx.x >> 1;  // 1: the Loewdin workgroup of a split call
  const DiagIso d = a.iso[iq];
  const int n = d.nst;
  if (a.ctl && a.ctl[iq] != 0u) return;  // the wide path has this block
  if (n <= 0) {
    if (a.construct && role == 0 && !a.ctl) diag_finish_arrive(a);
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

  // the wavefront Jacobi (blocks in LDS): matrix buffers in B2..B3
  // (m^2 <= 2 n^2 for n >= 2) and in B0 (even n) or after the rotation log
  // (odd n), V in B1, the log after B4
  double* const jlc = (double*)(B4 + n * n);
  cx<double>* const jls = (cx<double>*)(jlc + 2 * (jw_m(n) - 1) * (jw_m(n) / 2));
  cx<double>* const jA1 = (n & 1) ? (cx<double>*)((char*)jlc + jw_log_bytes(n)) : B0;
  cx<double>* Urho = B4;
  cx<double>* U = Urho;
  if (a.mode == DIAG_GIVEN_U) {
    // construct_hfb_matrices with the caller's unitary (static.py:484-577)
    tile_for<NP, NT>(n, [&](int i, int j) { Urho[i * n + j] = d.U[i * n + j]; });
    __syncthreads();
  } else {
    if (!a.split || role == 1) {
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
          jacobi_wave<NP, NT>(d.S, B2, jA1, B1, jlc, jls, n, jsh, false, B2);
        } else {
          load_lower_packed<NT>(Ap, d.S, n);
          jacobi_packed<NT>(Ap, Vt, B1, n, jsh, false);
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

      if (a.stats && tid == 0) {
        a.stats[2 + iq] = ok ? ns_iters : -1;
        if (!a.split) a.stats[iq] = 0;
      }
    }
    if (a.split && role == 1) {
      // hand U_rho over to the eigensolver's workgroup of this isospin: the
      // stores, then a release at device scope (the other workgroup may sit
      // on another XCD)
      tile_for<NP, NT>(n, [&](int i, int j) { d.U[i * n + j] = Urho[i * n + j]; });
      __syncthreads();
      if (tid == 0) __hip_atomic_store(&a.sync[1 + iq], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    // subspace diagonalisation (static.py:300-303): U <- U_rho U_lam
    if (a.tdiag && d.Lam) {
      int sw;
      if constexpr (WM == WM_LDS) {
        sw = jacobi_wave<NP, NT>(d.Lam, B2, jA1, B1, jlc, jls, n, jsh, true, B2, a.tol2);
      } else {
        load_lower_packed<NT>(Ap, d.Lam, n);
        sw = jacobi_packed<NT>(Ap, Vt, B1, n, jsh, true, a.tol2);
      }
      if (a.stats && tid == 0) a.stats[iq] = sw;
      if (a.split) {
        // U_rho from the Loewdin workgroup (acquire at device scope; the flag
        // returns to zero for the next call)
        if (tid == 0) {
          while (__hip_atomic_load(&a.sync[1 + iq], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0u)
            __builtin_amdgcn_s_sleep(2);
          __hip_atomic_store(&a.sync[1 + iq], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        tile_for<NP, NT>(n, [&](int i, int j) { Urho[i * n + j] = d.U[i * n + j]; });
        __syncthreads();
      }
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
  // with the wide path, its last kernel combines the isospins
  if (a.construct && !a.ctl) diag_finish_arrive(a);
}

// ---------------------------------------------------------------------------
// The Loewdin step and construct_hfb_matrices on many compute units, for
// blocks above NDIAG_LDS states without the subspace diagonalisation (the
// steady iterations of the large nuclei, static.py:295-350, 463-577).  One
// workgroup per isospin (diag_kernel) spends most of its time in the matrix
// products of these steps; here every product is a launch with one
// workgroup per 16 x 16 output tile, the four wavefronts of a workgroup
// taking a quarter of K each and adding their partial tiles through LDS.
// wd_prep_kernel forms E = S - I and chooses the series degree from |E|_F
// as diag_kernel does; a block that needs the Newton-Schulz iteration or
// the eigen-decomposition goes to diag_kernel, launched next with ctl
// telling it which blocks are its own.  Work matrices (dense n x n in the
// block's work area W): 0 E, then lambda+; 1 and 2 the Horner terms, then
// H_old U and G_old U, then (1) the lambda partial sums; 3 H; 4 G.
// ---------------------------------------------------------------------------
constexpr int WD_NT = 256;

__device__ __forceinline__ cx<double>* wd_slot(const DiagIso& d, int s) { return d.W + (size_t)s * d.nst * d.nst; }

__global__ void __launch_bounds__(1024) wd_prep_kernel(DiagArgs a) {
  __shared__ double red[32];
  const int iq = blockIdx.x;
  const DiagIso d = a.iso[iq];
  const int n = d.nst;
  if (n <= 0) {
    if (threadIdx.x == 0) a.ctl[iq] = 1u;
    return;
  }
  cx<double>* const E = wd_slot(d, 0);
  double fr = 0.0, dummy = 0.0;
  for (int e = threadIdx.x; e < n * n; e += 1024) {
    const int i = e / n, j = e - i * n;
    cx<double> v;
    if (i > j)
      v = d.S[i * n + j];
    else if (i < j)
      v = cconj(d.S[j * n + i]);
    else
      v = {d.S[i * n + i].x - 1.0, 0.0};
    E[e] = v;
    fr += cabs2(v);
    if (i == j) a.sp_norm[d.start + i] = d.S[i * n + i].x;
  }
  block_sum2_all<1024>(fr, dummy, red);
  if (threadIdx.x == 0) {
    const unsigned deg = fr < 1e-6 ? 4u : (fr < 1e-4 ? 7u : 0u);
    a.ctl[iq] = deg;
    if (a.stats && deg != 0u) {
      a.stats[2 + iq] = 0;
      a.stats[iq] = 0;
    }
  }
}

// This wavefront's quarter of K of the 16 x 16 tile (ti, tj) of op(A) B
// (op(A) = A or A^H; leading dimensions lda, ldb; entries beyond n are
// zeros), in the D layout of v_mfma_f64_16x16x4f64; operands of four
// k-steps are loaded together ahead of their products.
__device__ __forceinline__ void wd_tile(const cx<double>* __restrict__ A, int lda, bool conjA,
                                        const cx<double>* __restrict__ B, int ldb, int n, int ti, int tj, mm_d4& cr,
                                        mm_d4& ci) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nstep = ((n + 15) >> 4);  // k-steps of 4 per wavefront: K padded to 16, split four ways
  const int k0 = w * 4 * nstep;
  const int i = ti * 16 + li, j = tj * 16 + li;
  cr = mm_d4{0.0, 0.0, 0.0, 0.0};
  ci = mm_d4{0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < nstep; s += 4) {
    cx<double> av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + 4 * (s + u) + lk;
      const bool kin = s + u < nstep && k < n;
      const bool ka = kin && i < n, kb = kin && j < n;
      const int kc = kin ? k : 0, ic = ka ? i : 0, jc = kb ? j : 0;
      const cx<double> x = conjA ? cconj(A[kc * lda + ic]) : A[ic * lda + kc];
      const cx<double> y = B[kc * ldb + jc];
      av[u] = ka ? x : cx<double>{0.0, 0.0};
      bv[u] = kb ? y : cx<double>{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (s + u < nstep) {
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, bv[u].x, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, bv[u].y, ci, 0, 0, 0);
        cr = __builtin_amdgcn_mfma_f64_16x16x4f64(-av[u].y, bv[u].y, cr, 0, 0, 0);
        ci = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].y, bv[u].x, ci, 0, 0, 0);
      }
    }
  }
}

// The four K quarters summed (in wavefront order) through LDS; wavefront w
// returns the entry (row ti*16 + lane/16 + 4w, column tj*16 + lane%16).
__device__ __forceinline__ cx<double> wd_sum(const mm_d4& cr, const mm_d4& ci, double (*red)[8][64]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    red[w][r][lane] = cr[r];
    red[w][4 + r][lane] = ci[r];
  }
  __syncthreads();
  cx<double> v = {red[0][w][lane], red[0][4 + w][lane]};
#pragma unroll
  for (int q = 1; q < 4; ++q) {
    v.x += red[q][w][lane];
    v.y += red[q][4 + w][lane];
  }
  return v;
}

enum { WD_HORNER = 0, WD_HU = 1, WD_UHB = 2, WD_M2 = 3 };

// One product of the wide path per launch, blockIdx = (tile, isospin,
// right-hand block: 0 H, 1 G).
//   WD_HORNER, step j of the series of degree deg (static.py:463-482,
//     (I + E)^{-1/2} = sum_k c_k E^k by Horner from k = deg down to 0):
//     j = 0: T = c_deg E E + c_{deg-1} E + c_{deg-2} I, then
//     T <- E T + c_k I with k = deg - 2 - j; the k = 0 term is U (d.U).
//     Steps beyond the block's degree return at once.
//   WD_HU:  H_old U (z = 0), G_old U (z = 1) into work matrices 1, 2.
//   WD_UHB: H = U^H (H_old U), G likewise into work matrices 3, 4.
//   WD_M2:  M2 = U lambda+ (lambda+ in work matrix 0, or the stored lambda
//     when construct is off, static.py:336-350, 557-570); the extra last
//     workgroup of isospin 0 combines the fluctuation measures.
template <int OP>
__global__ void __launch_bounds__(WD_NT) wd_prod_kernel(DiagArgs a, int j) {
  __shared__ double red[4][8][64];
  const int iq = blockIdx.y;
  const DiagIso d = a.iso[iq];
  const int n = d.nst;
  if (OP == WD_M2 && blockIdx.x == gridDim.x - 1) {
    if (iq == 0 && a.construct && threadIdx.x < 64) {
      // efluct1q/2q of the wide blocks from wd_lambda_kernel's partial
      // sums, then efluct1 = max, efluct2 = mean over the isospins
      // (static.py:571-576)
      const int lane = threadIdx.x;
      for (int q = 0; q < 2; ++q) {
        const DiagIso dq = a.iso[q];
        const int nq = dq.nst;
        if (nq <= 0 || a.ctl[q] < 4u) continue;
        const int nwg = (nq * nq + WD_NT - 1) / WD_NT;
        const double* part = (const double*)wd_slot(dq, 1);
        double mx = 0.0, sm = 0.0;
        for (int b = lane; b < nwg; b += 64) {
          mx = fmax(mx, part[2 * b]);
          sm += part[2 * b + 1];
        }
        for (int o = 32; o > 0; o >>= 1) {
          mx = fmax(mx, __shfl_xor(mx, o, 64));
          sm += __shfl_xor(sm, o, 64);
        }
        if (lane == 0) {
          a.efl1q[q] = mx;
          a.efl2q[q] = sqrt(sm / ((double)nq * nq));
        }
      }
      if (lane == 0) {
        a.efl1[0] = fmax(a.efl1q[0], a.efl1q[1]);
        a.efl2[0] = 0.5 * (a.efl2q[0] + a.efl2q[1]);
      }
    }
    return;
  }
  const unsigned deg = a.ctl[iq];
  if (n <= 0 || deg < 4u) return;
  const int nt = (n + 15) >> 4;
  if ((int)blockIdx.x >= nt * nt) return;
  const int ti = blockIdx.x / nt, tj = blockIdx.x - ti * nt;
  const cx<double>* A;
  const cx<double>* B;
  int lda = n, ldb = n;
  bool conjA = false;
  cx<double>* C;
  double alpha = 1.0, beta = 0.0, gamma = 0.0;
  const cx<double>* Eadd = nullptr;
  if constexpr (OP == WD_HORNER) {
    constexpr double ck[8] = {1.0, -0.5, 0.375, -0.3125, 0.2734375, -0.24609375, 0.2255859375, -0.20947265625};
    const int k = (int)deg - 2 - j;
    if (k < 0) return;
    A = wd_slot(d, 0);
    if (j == 0) {
      B = A;
      Eadd = A;
      alpha = ck[deg];
      beta = ck[deg - 1];
    } else {
      B = wd_slot(d, 1 + ((j - 1) & 1));
    }
    gamma = ck[k];
    C = k == 0 ? d.U : wd_slot(d, 1 + (j & 1));
  } else if constexpr (OP == WD_HU) {
    if (blockIdx.z == 1 && !(a.ipair && d.Gold)) return;
    A = blockIdx.z == 0 ? d.Hold : d.Gold;
    B = d.U;
    C = wd_slot(d, 1 + blockIdx.z);
  } else if constexpr (OP == WD_UHB) {
    if (blockIdx.z == 1 && !(a.ipair && d.Gold)) return;
    A = d.U;
    conjA = true;
    B = wd_slot(d, 1 + blockIdx.z);
    C = wd_slot(d, 3 + blockIdx.z);
  } else {
    A = d.U;
    if (a.construct) {
      B = wd_slot(d, 0);
    } else {
      B = d.Ls;
      ldb = a.nmax;
    }
    C = d.M2;
  }
  mm_d4 cr, ci;
  wd_tile(A, lda, conjA, B, ldb, n, ti, tj, cr, ci);
  cx<double> v = wd_sum(cr, ci, red);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = ti * 16 + (lane >> 4) + 4 * w, col = tj * 16 + (lane & 15);
  if (row < n && col < n) {
    v = cscale(v, alpha);
    if (Eadd) v = cadd(v, cscale(Eadd[row * n + col], beta));
    if (row == col) v.x += gamma;
    C[row * n + col] = v;
  }
}

// H, G and lambda of construct_hfb_matrices elementwise (static.py:540-576):
// the static hmatrix/gapmatrix rows, sp_energy, deltaf, the off-diagonal
// gap entries zeroed with tbcs, lambda's anti-Hermitian part (symcond),
// its Hermitian part (lambda_save, and lambda+ into work matrix 0) and per
// workgroup the largest |anti-Hermitian entry| and the sum of their squares.
__global__ void __launch_bounds__(WD_NT) wd_lambda_kernel(DiagArgs a) {
  __shared__ double red[32];
  const int iq = blockIdx.y;
  const DiagIso d = a.iso[iq];
  const int n = d.nst;
  if (n <= 0 || a.ctl[iq] < 4u || (int)blockIdx.x * WD_NT >= n * n) return;
  const int nm = a.nmax;
  const bool gp = a.ipair && d.Gold;
  const cx<double>* H = wd_slot(d, 3);
  const cx<double>* G = wd_slot(d, 4);
  const int e = blockIdx.x * WD_NT + threadIdx.x;
  double mx = 0.0, sm = 0.0;
  if (e < n * n) {
    const int i = e / n, j = e - i * n;
    const bool zg = !gp || (a.tbcs && i != j);
    const cx<double> hij = H[i * n + j], hji = H[j * n + i];
    const cx<double> gij = zg ? cx<double>{0.0, 0.0} : G[i * n + j];
    const cx<double> gji = zg ? cx<double>{0.0, 0.0} : G[j * n + i];
    d.Hm[i * nm + j] = hij;
    d.Gm[i * nm + j] = gij;
    if (i == j) {
      a.sp_energy[d.start + i] = hij.x;
      if (a.ipair) a.deltaf[d.start + i] = gij.x * a.pairwg[d.start + i];
    }
    const int si = d.start + i, sj = d.start + j;
    const double wj = a.wocc[sj] * a.wstates[sj];
    const double uj = a.wguv[sj] * a.pairwg[sj] * a.wstates[sj];
    const double wi = a.wocc[si] * a.wstates[si];
    const double ui = a.wguv[si] * a.pairwg[si] * a.wstates[si];
    const cx<double> lij = csub(cscale(hij, wj), cscale(gij, uj));
    const cx<double> lji = csub(cscale(hji, wi), cscale(gji, ui));
    const cx<double> lh = cconj(lji);
    const cx<double> asym = cscale(csub(lij, lh), 0.5);
    const cx<double> sym = cscale(cadd(lij, lh), 0.5);
    d.Sym[i * nm + j] = asym;
    d.Ls[i * nm + j] = sym;
    wd_slot(d, 0)[i * n + j] = sym;
    const double aa = cabs2(asym);
    mx = sqrt(aa);
    sm = aa;
  }
  const double M = block_max_all<WD_NT>(mx, red);
  double S = sm, dummy = 0.0;
  block_sum2_all<WD_NT>(S, dummy, red);
  if (threadIdx.x == 0) {
    double* part = (double*)wd_slot(d, 1);
    part[2 * blockIdx.x] = M;
    part[2 * blockIdx.x + 1] = S;
  }
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
// the instance (LDS up to NDIAG_LDS states, global memory above).  With
// a.ctl the wide kernels surround it: wd_prep_kernel before (it decides
// which blocks diag_kernel keeps), the products after.
static int launch_diag(const DiagArgs& a, int nbig, hipStream_t st) {
#define HFB_DIAG(NPV, NTV, WMV, LDSV)                                                        \
  {                                                                                         \
    if ((LDSV) > 32 * 1024) allow_full_lds((const void*)diag_kernel<NPV, NTV, WMV>);        \
    hipLaunchKernelGGL((diag_kernel<NPV, NTV, WMV>), dim3(b.split ? 4 : 2), dim3(NTV), (LDSV), st, b); \
  }
  // blocks in LDS with the subspace diagonalisation: Loewdin and the
  // eigensolver on separate workgroups (four in all, each its own compute
  // unit: the eigensolver's workgroup waits for its Loewdin workgroup)
  DiagArgs b = a;
  b.split = nbig <= NDIAG_LDS && a.tdiag && a.mode == DIAG_FULL;
  if (a.ctl) {
    hipLaunchKernelGGL(wd_prep_kernel, dim3(2), dim3(1024), 0, st, a);
    const int st0 = launch_status();
    if (st0 != HFB_OK) return st0;
  }
  const size_t cz = sizeof(cx<double>);
  const size_t tri = (size_t)nbig * (nbig + 1) / 2;
  if (nbig <= NDIAG_LDS) {
    // each block's own layout; the allocation covers both
    size_t lds = 0;
    for (int iq = 0; iq < 2; ++iq) {
      const int nq = a.iso[iq].nst;
      if (nq > 0) lds = std::max(lds, cz * 5 * (size_t)nq * nq + jw_extra_bytes(nq));
    }
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
  if (a.ctl) {
    const int nt = (nbig + 15) / 16, tiles = nt * nt;
    const int nz = a.ipair ? 2 : 1;
    for (int j = 0; j < 6; ++j)
      hipLaunchKernelGGL(wd_prod_kernel<WD_HORNER>, dim3(tiles, 2, 1), dim3(WD_NT), 0, st, a, j);
    if (a.construct) {
      hipLaunchKernelGGL(wd_prod_kernel<WD_HU>, dim3(tiles, 2, nz), dim3(WD_NT), 0, st, a, 0);
      hipLaunchKernelGGL(wd_prod_kernel<WD_UHB>, dim3(tiles, 2, nz), dim3(WD_NT), 0, st, a, 0);
      hipLaunchKernelGGL(wd_lambda_kernel, dim3((nbig * nbig + WD_NT - 1) / WD_NT, 2), dim3(WD_NT), 0, st, a);
    }
    hipLaunchKernelGGL(wd_prod_kernel<WD_M2>, dim3(tiles + 1, 2, 1), dim3(WD_NT), 0, st, a, 0);
  }
  return launch_status();
}

// mats[iq*4 + m]: m=0 S, 1 H_old, 2 lambda_lin (or null), 3 G_old (or null),
// each nst_iq x nst_iq complex; uo[iq*2]: U, uo[iq*2+1]: M2 outputs.
extern "C" int hfb_diag_small(hfb_ctx* c, int nst0, int nst1, int nmax, const double* const* mats, int tdiag, int ipair,
                              int tbcs, int construct, const double* wocc, const double* wstates, const double* wguv,
                              const double* pairwg, double* sp_norm, double* sp_energy, double* deltaf,
                              double* hmatrix, double* gapmatrix, double* symcond, double* lambda_save,
                              double* efluct1q, double* efluct2q, double* efluct1, double* efluct2,
                              double* const* uo, double* work, double* stats, double jacobi_tol,
                              void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!(jacobi_tol >= 0.0) || jacobi_tol >= 1.0) return HFB_E_ARG;
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
  a.tol2 = jacobi_tol > 0.0 ? jacobi_tol * jacobi_tol : JACOBI_TOL2;
  a.ipair = ipair;
  a.tbcs = tbcs;
  a.construct = construct;
  a.stats = stats;
  // blocks above NDIAG_LDS states without the diagonalisation: the wide path
  if (nbig0 > NDIAG_LDS && !tdiag) a.ctl = c->sync + HFB_SYNC_DIAG_CTL;
  return launch_diag(a, nbig0, (hipStream_t)stream);
}

// Loewdin S^{-1/2} of one block (static.py:463-482): U (nst x nst complex)
// and sp_norm = Re diag S; work: null up to NDIAG_LDS states, else 5 nst^2
// complex doubles.
extern "C" int hfb_loewdin(int nst, const double* S, double* U, double* sp_norm, double* work, void* stream) {
  if (!S || !U || !sp_norm) return HFB_E_ARG;
  if (nst <= 0 || nst > NDIAG_MAX) return HFB_E_SIZE;
  if (nst > NDIAG_LDS && !work) return HFB_E_ARG;
  DiagArgs a{};
  a.tol2 = JACOBI_TOL2;
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
  a.tol2 = JACOBI_TOL2;
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
