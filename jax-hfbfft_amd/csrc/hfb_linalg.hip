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

constexpr int NMAXJ = 64;

struct DiagIso {
  int nst, start;
  const cx<double>* S;
  const cx<double>* Hold;
  const cx<double>* Lam;   // null unless diagonalising
  const cx<double>* Gold;  // null unless pairing
  cx<double>* U;
  cx<double>* M2;
  cx<double>* T1;  // scratch nst x nst (used when the work matrices do not fit in LDS)
  cx<double>* T2;
  cx<double>* T3;
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
  int tdiag, ipair, tbcs, construct;
};

__device__ __forceinline__ cx<double> cmulc(cx<double> a, cx<double> b) { return cmul(a, b); }

// A (n x n, LDS) <- Hermitian matrix from the lower triangle of M (global).
__device__ void load_lower_herm(cx<double>* A, const cx<double>* M, int n) {
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    cx<double> v;
    if (i > j)
      v = M[i * n + j];
    else if (i < j)
      v = cconj(M[j * n + i]);
    else
      v = {M[i * n + i].x, 0.0};
    A[e] = v;
  }
  __syncthreads();
}

struct JacobiShared {
  int p[NMAXJ / 2], q[NMAXJ / 2];
  double c[NMAXJ / 2], s[NMAXJ / 2];
  cx<double> e[NMAXJ / 2];
  double red[32];
  int perm[NMAXJ];
  double ev[NMAXJ];
  int done;
};

// Eigen-decomposition of the Hermitian A (LDS, destroyed); V (LDS) gets the
// eigenvectors as columns, js.ev the eigenvalues; sorted ascending when
// `sort` (T is global scratch used for the column permutation).
__device__ void jacobi_eigh(cx<double>* A, cx<double>* V, int n, JacobiShared& js, bool sort, cx<double>* T) {
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += blockDim.x) V[e] = cx<double>{(e / n == e % n) ? 1.0 : 0.0, 0.0};
  const int m = n + (n & 1);
  const int npair = m / 2;
  __syncthreads();
  for (int sweep = 0; sweep < 40; ++sweep) {
    double v[2] = {0.0, 0.0};
    for (int e = tid; e < n * n; e += blockDim.x) {
      const int i = e / n, j = e % n;
      if (i == j)
        v[1] += cabs2(A[e]);
      else
        v[0] += cabs2(A[e]);
    }
    block_sum<2>(v, js.red);
    if (tid == 0) js.done = (v[0] <= 1e-30 * v[1]) || v[0] == 0.0;
    __syncthreads();
    if (js.done) break;
    for (int t = 0; t < m - 1; ++t) {
      if (tid < npair) {
        const int k = tid;
        const int a0 = (k == 0) ? 0 : ((k - 1 + t) % (m - 1)) + 1;
        const int kb = m - 1 - k;
        const int b0 = ((kb - 1 + t) % (m - 1)) + 1;
        int p = a0 < b0 ? a0 : b0, q = a0 < b0 ? b0 : a0;
        double c = 1.0, s = 0.0;
        cx<double> ea = {1.0, 0.0};
        if (q < n) {
          const double app = A[p * n + p].x, aqq = A[q * n + q].x;
          const cx<double> apq = A[p * n + q];
          const double r = sqrt(cabs2(apq));
          if (r > 1e-300 && r > 1e-18 * (fabs(app) + fabs(aqq))) {
            ea = {apq.x / r, apq.y / r};
            const double th = (aqq - app) / (2.0 * r);
            const double tt = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(1.0 + th * th));
            c = 1.0 / sqrt(1.0 + tt * tt);
            s = tt * c;
          }
        } else {
          p = q = -1;
        }
        js.p[k] = p;
        js.q[k] = q;
        js.c[k] = c;
        js.s[k] = s;
        js.e[k] = ea;
      }
      __syncthreads();
      // columns of A and V: X <- X V2
      for (int it = tid; it < 2 * npair * n; it += blockDim.x) {
        const int mat = it / (npair * n);
        const int rem = it % (npair * n);
        const int k = rem / n, r = rem % n;
        const int p = js.p[k], q = js.q[k];
        if (p < 0 || js.s[k] == 0.0) continue;
        cx<double>* X = mat ? V : A;
        const cx<double> xp = X[r * n + p], xq = X[r * n + q];
        const cx<double> ec = cscale(js.e[k], js.c[k]), es = cscale(js.e[k], js.s[k]);
        X[r * n + p] = csub(cmul(xp, ec), cscale(xq, js.s[k]));
        X[r * n + q] = cadd(cmul(xp, es), cscale(xq, js.c[k]));
      }
      __syncthreads();
      // rows of A: A <- V2^H A
      for (int it = tid; it < npair * n; it += blockDim.x) {
        const int k = it / n, r = it % n;
        const int p = js.p[k], q = js.q[k];
        if (p < 0 || js.s[k] == 0.0) continue;
        const cx<double> xp = A[p * n + r], xq = A[q * n + r];
        const cx<double> ecc = cscale(cconj(js.e[k]), js.c[k]), esc = cscale(cconj(js.e[k]), js.s[k]);
        A[p * n + r] = csub(cmul(ecc, xp), cscale(xq, js.s[k]));
        A[q * n + r] = cadd(cmul(esc, xp), cscale(xq, js.c[k]));
      }
      __syncthreads();
      if (tid < npair) {
        const int p = js.p[tid], q = js.q[tid];
        if (p >= 0 && js.s[tid] != 0.0) {
          A[p * n + q] = {0.0, 0.0};
          A[q * n + p] = {0.0, 0.0};
          A[p * n + p].y = 0.0;
          A[q * n + q].y = 0.0;
        }
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < n; i += blockDim.x) js.ev[i] = A[i * n + i].x;
  __syncthreads();
  if (sort) {
    if (tid == 0) {
      for (int i = 0; i < n; ++i) js.perm[i] = i;
      for (int i = 0; i < n; ++i) {  // stable selection sort
        int b = i;
        for (int j = i + 1; j < n; ++j)
          if (js.ev[js.perm[j]] < js.ev[js.perm[b]]) b = j;
        const int tmp = js.perm[b];
        for (int j = b; j > i; --j) js.perm[j] = js.perm[j - 1];
        js.perm[i] = tmp;
      }
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += blockDim.x) T[e] = V[(e / n) * n + js.perm[e % n]];
    __syncthreads();
    for (int e = tid; e < n * n; e += blockDim.x) V[e] = T[e];
    double evs = 0.0;
    if (tid < n) evs = js.ev[js.perm[tid]];
    __syncthreads();
    if (tid < n) js.ev[tid] = evs;
    __syncthreads();
  }
}

// C = op(A) B (n x n); opA 0: A, 1: A^H.  Strides in elements.
__device__ void mm(cx<double>* C, int ldc, const cx<double>* A, int lda, int opA, const cx<double>* B, int ldb, int n) {
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e % n;
    cx<double> acc = {0.0, 0.0};
    for (int k = 0; k < n; ++k) {
      const cx<double> a = opA ? cconj(A[k * lda + i]) : A[i * lda + k];
      acc = cadd(acc, cmul(a, B[k * ldb + j]));
    }
    C[i * ldc + j] = acc;
  }
  __syncthreads();
}

constexpr int NLDS_ALL = 40;  // up to this many states all five work matrices sit in LDS

// Inverse square root of the Hermitian positive matrix in Y (scaled copy of
// S) by the coupled Newton-Schulz iteration Y <- Y T, Z <- T Z,
// T = (3I - Z Y)/2, which converges to Z = Y0^{-1/2}.  Returns false when it
// has not converged (nearly singular S); the caller then takes the
// eigen-decomposition with the reference's clamp.
__device__ bool newton_schulz(cx<double>*& Y, cx<double>*& Z, cx<double>* P, cx<double>*& Y2, cx<double>*& Z2, int n,
                              double* red) {
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += blockDim.x) Z[e] = cx<double>{(e / n == e % n) ? 1.0 : 0.0, 0.0};
  __syncthreads();
  __shared__ double resid;
  int remaining = -1;  // updates left once the residual is below 1e-6
  for (int it = 0; it < 80; ++it) {
    mm(P, n, Z, n, 0, Y, n, n);
    double v = 0.0;
    for (int e = tid; e < n * n; e += blockDim.x) {
      const cx<double> d = csub(P[e], cx<double>{(e / n == e % n) ? 1.0 : 0.0, 0.0});
      v = fmax(v, fabs(d.x) + fabs(d.y));
    }
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      double m = 0.0;
      for (int w = 0; w < (int)blockDim.x / 64; ++w) m = fmax(m, red[w]);
      resid = m;
    }
    __syncthreads();
    // quadratic convergence: two updates after 1e-6 reach round-off
    if (remaining < 0 && resid < 1e-6) remaining = 2;
    for (int e = tid; e < n * n; e += blockDim.x) {  // T = (3I - P)/2
      const double id = (e / n == e % n) ? 1.5 : 0.0;
      P[e] = cx<double>{id - 0.5 * P[e].x, -0.5 * P[e].y};
    }
    __syncthreads();
    mm(Y2, n, Y, n, 0, P, n, n);
    mm(Z2, n, P, n, 0, Z, n, n);
    cx<double>* t = Y;
    Y = Y2;
    Y2 = t;
    t = Z;
    Z = Z2;
    Z2 = t;
    if (remaining > 0 && --remaining == 0) return true;
  }
  return false;
}

__global__ void __launch_bounds__(256) diag_kernel(DiagArgs a) {
  extern __shared__ double lds_raw[];
  __shared__ JacobiShared js;
  __shared__ double red[8];
  const int iq = blockIdx.x;
  const DiagIso d = a.iso[iq];
  const int n = d.nst;
  if (n <= 0) return;
  const int tid = threadIdx.x;
  const int nm = a.nmax;
  cx<double>* Bf[5];
  Bf[0] = (cx<double>*)lds_raw;
  Bf[1] = Bf[0] + n * n;
  if (n <= NLDS_ALL) {
    Bf[2] = Bf[1] + n * n;
    Bf[3] = Bf[2] + n * n;
    Bf[4] = Bf[3] + n * n;
  } else {
    Bf[2] = d.T1;
    Bf[3] = d.T2;
    Bf[4] = d.T3;
  }

  // Loewdin S^{-1/2} (static.py:463-482)
  for (int i = tid; i < n; i += blockDim.x) a.sp_norm[d.start + i] = d.S[i * n + i].x;
  load_lower_herm(Bf[0], d.S, n);
  __shared__ double cscale_s;
  if (tid == 0) {  // Gershgorin bound on the largest eigenvalue
    double c = 0.0;
    for (int i = 0; i < n; ++i) {
      double r = 0.0;
      for (int j = 0; j < n; ++j) r += sqrt(cabs2(Bf[0][i * n + j]));
      c = fmax(c, r);
    }
    cscale_s = c > 0.0 ? c : 1.0;
  }
  __syncthreads();
  const double cs = cscale_s;
  for (int e = tid; e < n * n; e += blockDim.x) Bf[0][e] = cscale(Bf[0][e], 1.0 / cs);
  __syncthreads();
  cx<double>*Y = Bf[0], *Z = Bf[1], *Y2 = Bf[3], *Z2 = Bf[4];
  const bool ok = newton_schulz(Y, Z, Bf[2], Y2, Z2, n, red);
  cx<double>* Urho = Bf[4];
  if (ok) {
    const double f = 1.0 / sqrt(cs);
    // Z may live in Bf[4]; scale in place or copy
    for (int e = tid; e < n * n; e += blockDim.x) {
      const cx<double> v = cscale(Z[e], f);
      Urho[e] = v;
    }
    __syncthreads();
  } else {
    load_lower_herm(Bf[0], d.S, n);
    jacobi_eigh(Bf[0], Bf[1], n, js, false, Bf[2]);
    __shared__ double fac[NMAXJ];
    if (tid == 0) {
      double wmax = js.ev[0];
      for (int i = 1; i < n; ++i) wmax = fmax(wmax, js.ev[i]);
      const double thr = fmax(1e-12, 2.220446049250313e-16 * wmax * n);
      for (int i = 0; i < n; ++i) fac[i] = 1.0 / sqrt(fmax(js.ev[i], thr));
    }
    __syncthreads();
    const cx<double>* V = Bf[1];
    for (int e = tid; e < n * n; e += blockDim.x) {  // U = V diag(fac) V^H
      const int i = e / n, j = e % n;
      cx<double> acc = {0.0, 0.0};
      for (int k = 0; k < n; ++k) acc = cadd(acc, cscale(cmul(V[i * n + k], cconj(V[j * n + k])), fac[k]));
      Urho[e] = acc;
    }
    __syncthreads();
  }

  // subspace diagonalisation (static.py:300-303): U <- U_rho U_lam
  cx<double>* U = Urho;
  if (a.tdiag && d.Lam) {
    load_lower_herm(Bf[0], d.Lam, n);
    jacobi_eigh(Bf[0], Bf[1], n, js, true, Bf[2]);
    mm(Bf[3], n, Urho, n, 0, Bf[1], n, n);
    U = Bf[3];
  }
  for (int e = tid; e < n * n; e += blockDim.x) d.U[e] = U[e];

  if (a.construct) {  // construct_hfb_matrices (static.py:484-577)
    cx<double>* H = Bf[2];
    cx<double>* G = Bf[0];
    for (int e = tid; e < n * n; e += blockDim.x) Bf[0][e] = d.Hold[e];
    __syncthreads();
    mm(Bf[1], n, Bf[0], n, 0, U, n, n);
    mm(H, n, U, n, 1, Bf[1], n, n);
    if (a.ipair && d.Gold) {
      for (int e = tid; e < n * n; e += blockDim.x) Bf[0][e] = d.Gold[e];
      __syncthreads();
      mm(Bf[1], n, Bf[0], n, 0, U, n, n);
      mm(G, n, U, n, 1, Bf[1], n, n);
    } else {
      for (int e = tid; e < n * n; e += blockDim.x) G[e] = {0.0, 0.0};
      __syncthreads();
    }
    for (int e = tid; e < n * n; e += blockDim.x) {
      d.Hm[(e / n) * nm + e % n] = H[e];
      d.Gm[(e / n) * nm + e % n] = G[e];
    }
    for (int i = tid; i < n; i += blockDim.x) {
      a.sp_energy[d.start + i] = H[i * n + i].x;
      if (a.ipair) a.deltaf[d.start + i] = G[i * n + i].x * a.pairwg[d.start + i];
    }
    __syncthreads();
    if (a.tbcs) {
      for (int e = tid; e < n * n; e += blockDim.x)
        if (e / n != e % n) {
          G[e] = {0.0, 0.0};
          d.Gm[(e / n) * nm + e % n] = {0.0, 0.0};
        }
      __syncthreads();
    }
    // lambda, its parts and the fluctuation measures; lambda+ kept in Bf[1]
    double mx = 0.0, sm = 0.0;
    for (int e = tid; e < n * n; e += blockDim.x) {
      const int i = e / n, j = e % n;
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
      Bf[1][e] = sym;
      const double aa = cabs2(asym);
      mx = fmax(mx, sqrt(aa));
      sm += aa;
    }
    __shared__ double rmax[16], rsum[16];
    for (int o = 32; o > 0; o >>= 1) {
      mx = fmax(mx, __shfl_down(mx, o, 64));
      sm += __shfl_down(sm, o, 64);
    }
    if ((tid & 63) == 0) {
      rmax[tid >> 6] = mx;
      rsum[tid >> 6] = sm;
    }
    __syncthreads();
    if (tid == 0) {
      double M = 0.0, Ssum = 0.0;
      for (int w = 0; w < (int)blockDim.x / 64; ++w) {
        M = fmax(M, rmax[w]);
        Ssum += rsum[w];
      }
      a.efl1q[iq] = M;
      a.efl2q[iq] = sqrt(Ssum / ((double)n * n));
    }
    __syncthreads();
  } else {
    // construct=False: lagrange from the stored lambda (static.py:336-350)
    for (int e = tid; e < n * n; e += blockDim.x) Bf[1][e] = d.Ls[(e / n) * nm + e % n];
    __syncthreads();
  }
  // M2 = U lambda+ (static.py:557-570)
  mm(Bf[0], n, U, n, 0, Bf[1], n, n);
  for (int e = tid; e < n * n; e += blockDim.x) d.M2[e] = Bf[0][e];
}

__global__ void diag_finish_kernel(const double* efl1q, const double* efl2q, double* efl1, double* efl2) {
  efl1[0] = fmax(efl1q[0], efl1q[1]);
  efl2[0] = 0.5 * (efl2q[0] + efl2q[1]);
}

}  // namespace

// mats[iq*4 + m]: m=0 S, 1 H_old, 2 lambda_lin (or null), 3 G_old (or null),
// each nst_iq x nst_iq complex; uo[iq*2]: U, uo[iq*2+1]: M2 outputs;
// scratch: 6 * nmax^2 complex doubles.
extern "C" int hfb_diag_small(int nst0, int nst1, int nmax, const double* const* mats, int tdiag, int ipair,
                              int tbcs, int construct, const double* wocc, const double* wstates, const double* wguv,
                              const double* pairwg, double* sp_norm, double* sp_energy, double* deltaf,
                              double* hmatrix, double* gapmatrix, double* symcond, double* lambda_save,
                              double* efluct1q, double* efluct2q, double* efluct1, double* efluct2,
                              double* const* uo, double* scratch, void* stream) {
  if (!mats || !uo || !scratch || !wocc || !wstates || !wguv || !pairwg || !sp_norm || !sp_energy || !deltaf ||
      !hmatrix || !gapmatrix || !symcond || !lambda_save || !efluct1q || !efluct2q || !efluct1 || !efluct2)
    return HFB_E_ARG;
  if (nst0 < 0 || nst1 < 0 || nst0 > NMAXJ || nst1 > NMAXJ || nmax < nst0 || nmax < nst1) return HFB_E_SIZE;
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
    d.T1 = (cx<double>*)scratch + (iq * 3 + 0) * nm2;
    d.T2 = (cx<double>*)scratch + (iq * 3 + 1) * nm2;
    d.T3 = (cx<double>*)scratch + (iq * 3 + 2) * nm2;
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
  a.tdiag = tdiag;
  a.ipair = ipair;
  a.tbcs = tbcs;
  a.construct = construct;
  const int nbig = nst0 > nst1 ? nst0 : nst1;
  const size_t lds = sizeof(cx<double>) * (nbig <= NLDS_ALL ? 5 : 2) * (size_t)nbig * nbig;
  hipStream_t st = (hipStream_t)stream;
  if (lds > 32 * 1024) allow_full_lds((const void*)diag_kernel);
  // one wavefront per isospin for small blocks: the Jacobi sweeps are
  // barrier-bound and a single-wave barrier is nearly free
  const int nthr = nbig <= 32 ? 64 : 256;
  hipLaunchKernelGGL(diag_kernel, dim3(2), dim3(nthr), lds, st, a);
  int rc = launch_status();
  if (rc != HFB_OK) return rc;
  if (construct) {
    hipLaunchKernelGGL(diag_finish_kernel, dim3(1), dim3(1), 0, st, efluct1q, efluct2q, efluct1, efluct2);
    rc = launch_status();
  }
  return rc;
}
