/*
 * C ABI of libhfbhip.so -- the MI355X (gfx950) kernels of the static
 * Skyrme-HF(+BCS) iteration of ascsn/jax-hfbfft.
 *
 * Conventions
 *  - Every entry returns int: HFB_OK (0) or a negative HFB_E_* code;
 *    hfb_strerror() names it.  No C++ exception crosses this boundary.
 *  - The caller owns every array.  Arrays are device pointers in the
 *    reference's layouts: spinor blocks (S, 2, nx, ny, nz) complex, z fastest
 *    (levels.py:62-65); isospin fields (2, nx, ny, nz) and vector fields
 *    (2, 3, nx, ny, nz) (meanfield.py:24-25, densities.py:18-19).
 *  - prec selects the arithmetic of the spinor kernels: HFB_F64 (complex128
 *    arrays, double potentials) or HFB_F32 (complex64 arrays, float
 *    potentials).  Reductions accumulate in double.
 *  - stream is a hipStream_t (0 = default stream).  Every call is stream
 *    ordered and returns without synchronising unless stated.
 *  - The context owns only small tables (twiddles, derivative matrices,
 *    the Coulomb kernel); no entry allocates device memory after
 *    hfb_ctx_create.
 */
#ifndef HFB_HIP_H
#define HFB_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  HFB_OK = 0,
  HFB_E_ARG = -1,     /* bad argument (null pointer, bad enum, bad count) */
  HFB_E_SIZE = -2,    /* grid length without a compiled kernel */
  HFB_E_HIP = -3,     /* a HIP runtime call or launch failed */
  HFB_E_NOCTX = -4    /* null context */
};

enum { HFB_F64 = 0, HFB_F32 = 1 };

typedef struct hfb_ctx hfb_ctx;

/* Skyrme couplings and switches passed by value to the potential and
 * energy kernels; field meanings as in forces.py:7-58 (the b- and
 * C-couplings of forces.py:95-122), with h2m already c.m.-scaled as
 * init_static leaves it (static.py:87-88). */
typedef struct hfb_force {
  double b0, b0p, b1, b1p, b2, b2p, b3, b3p, b4, b4p, power, slate;
  double h2m[2];
  double Crho0, Crho1, Crho0D, Crho1D, Cdrho0, Cdrho1, Ctau0, Ctau1, CdJ0, CdJ1;
  double v0neut, v0prot, rho0pr;
  double e2, mass_number;
  int ipair, ex, zpe, tcoul;
} hfb_force;

/* Arguments of hfb_energy_tail as one record (hfb_weights_job.tail). */
typedef struct hfb_tail_job {
  const hfb_force* force;
  const double *sky_part, *en_part, *current;
  int nstates;
  const double *wocc, *wstates, *sp_kinetic, *sp_energy, *sp_orbital, *sp_spin;
  double* E;
} hfb_tail_job;

/* Arguments of hfb_grstep_weights as one record, for the launches that run
 * it in an extra workgroup (hfb_hpsi_ex).  tail (host memory, read at the
 * call; NULL for none): an hfb_energy_tail that the same h.psi call runs in
 * an extra workgroup of one of its launches; the weights then take the
 * v_pair maxima from the tail's skyrme partials (the values it writes to
 * E[HFB_EN_VMAX]) instead of vpair_max. */
typedef struct hfb_weights_job {
  int nstates, npsi0, tbcs;
  const double *wocc, *wstates, *wguv, *pairwg, *v_pair;
  long long nvox;
  const double* vpair_max;
  double *w, *wuv, *pw, *pshift;
  const struct hfb_tail_job* tail;
} hfb_weights_job;

/* Arguments of hfb_grstep_finish as one record (hfb_gram2_ex). */
typedef struct hfb_finish_job {
  int nstates;
  const double* hsum;
  int hsum_ntile;
  double wxyz;
  double* sp_energy;
  const double *wocc, *wstates, *fsum;
  int fsum_ntile;
  double *sp_efluct1, *sp_efluct2, *acc;
} hfb_finish_job;

/* Slots of the device energy vector (double[HFB_EN_LEN]), energies.py:6-43. */
enum {
  HFB_EN_EHFT = 0, HFB_EN_EHF0, HFB_EN_EHF1, HFB_EN_EHF2, HFB_EN_EHF3, HFB_EN_EHFLS, HFB_EN_EHFLSODD,
  HFB_EN_EHFC, HFB_EN_ECORC, HFB_EN_EHFINT, HFB_EN_E3CORR, HFB_EN_EZPE, HFB_EN_CRHO0, HFB_EN_CRHO1,
  HFB_EN_CDRHO0, HFB_EN_CDRHO1, HFB_EN_CTAU0, HFB_EN_CTAU1, HFB_EN_CDJ0, HFB_EN_CDJ1, HFB_EN_CJ0,
  HFB_EN_CJ1, HFB_EN_EHF, HFB_EN_TKE, HFB_EN_ORB /* 3 */, HFB_EN_SPIN = HFB_EN_ORB + 3,
  HFB_EN_JTOT = HFB_EN_SPIN + 3, HFB_EN_EFL1 = HFB_EN_JTOT + 3, HFB_EN_EFL2, HFB_EN_EFL1Q /* 2 */,
  HFB_EN_EFL2Q = HFB_EN_EFL1Q + 2, HFB_EN_ECORRP = HFB_EN_EFL2Q + 2, HFB_EN_SUMFLU, HFB_EN_DELESUM,
  HFB_EN_EPAIR /* 2 */, HFB_EN_JACOBI = HFB_EN_EPAIR + 2 /* 2 */, HFB_EN_NSITER = HFB_EN_JACOBI + 2 /* 2 */,
  HFB_EN_VMAX = HFB_EN_NSITER + 2 /* 2: max of v_pair per isospin, from hfb_skyrme */,
  HFB_EN_LEN = 50
};

const char* hfb_strerror(int code);

/* Grid + tables.  Replaces the device-side state of init_grids
 * (grids.py:66-138: coordinates, sder/sder2 matrices) and init_coulomb
 * (coulomb.py:23-52: FFT of the 1/r kernel on the doubled grid).  Each axis
 * length must be one of 24, 32, 48. */
int hfb_ctx_create(hfb_ctx** out, int device, int nx, int ny, int nz, double dx, double dy,
                   double dz, int periodic);
int hfb_ctx_destroy(hfb_ctx* ctx);

/* One-axis spectral derivatives of nfields fields (nx,ny,nz), complex.
 * kind 0: out1 = d/da                    (levels.py:131-189, cderv{x,y,z}00)
 * kind 1: out1 = d/da, out2 = d2/da2     (levels.py:190-281, cderv{x,y,z}01)
 * kind 2: out1 = d/da, out2 = d/da(B d/da) with the Nyquist rule
 *                                         (levels.py:283-395, cderv{x,y,z}02)
 * bmass: real (nx,ny,nz) field B for kind 2 (prec's real type). */
int hfb_cderv(hfb_ctx* ctx, int prec, int axis, int kind, const void* in, const void* bmass,
              void* out1, void* out2, int nfields, void* stream);

/* Single-particle Hamiltonian on nstates two-spinors (meanfield.py:40-159,
 * hpsi01), batched: state s uses the potentials of isospin (s >= npsi0).
 * Potentials: upot, bmass, v_pair (2,nx,ny,nz); aq, spot, wlspot
 * (2,3,nx,ny,nz), all in prec's real type.  w, wuv: per-state weights
 * (double).  Outputs (any may be null except pout_mf):
 *   pout    = w*h psi - wuv*v_pair psi,  pout_mf = h psi,  pout_del = v_pair psi.
 * partial (optional, double): per state and tile the pair
 *   (sum Re conj(psi) h psi, sum Re conj(psi) pout); reduce with
 *   hfb_reduce_partials.  ntile_out receives the tile count. */
int hfb_hpsi(hfb_ctx* ctx, int prec, const void* psi, int nstates, int npsi0, const void* upot,
             const void* bmass, const void* aq, const void* spot, const void* wlspot,
             const void* v_pair, const double* w, const double* wuv, void* pout, void* pout_mf,
             void* pout_del, double* partial, int* ntile_out, void* stream);

/* hfb_hpsi with hfb_grstep_weights (job, may be null) in one extra
 * workgroup of the first launch: the gradient step's weights
 * (static.py:152-170, 125-136) are written while the x pass runs, which
 * does not read them; the later passes and hfb_precond_step do.  job->tail
 * runs in one extra workgroup of the y-z plane launch (of the first launch
 * on grids without one); it writes only E, which no pass reads. */
int hfb_hpsi_ex(hfb_ctx* ctx, int prec, const void* psi, int nstates, int npsi0, const void* upot,
                const void* bmass, const void* aq, const void* spot, const void* wlspot, const void* v_pair,
                const double* w, const double* wuv, void* pout, void* pout_mf, void* pout_del, double* partial,
                int* ntile_out, const hfb_weights_job* job, void* stream);

/* Number of partial-sum tiles per state written by hfb_hpsi. */
int hfb_hpsi_ntiles(hfb_ctx* ctx, int prec);

/* out[s*nv + j] = scale * sum_t partial[(s*ntile + t)*nv + j]. */
int hfb_reduce_partials(const double* partial, int nstates, int ntile, int nv, double scale,
                        double* out, void* stream);

/* Self-test of the FP64 MFMA 16x16x4 lane layout used by the Gram kernels:
 * a (16x4) and b (4x16) row-major doubles in, d = a*b (16x16) out. */
int hfb_selftest_mfma_f64(const double* a, const double* b, double* d, void* stream);

/* Lane layout check of the FP32 16x16x4 matrix-core product the FP32 h.psi
 * kernels are built on (no counterpart in the reference): a (16, 4), b
 * (4, 16) float; d (64 lanes x 4 registers) float receives the raw result
 * registers, lane-major. */
int hfb_selftest_mfma_f32(const float* a, const float* b, float* d, void* stream);


/* ---------------------------------------------------------------------
 * Coulomb (coulomb.py:54-75, poisson)
 * --------------------------------------------------------------------- */

/* wcoul (nx,ny,nz) = Re IFFT3(e2 * wxyz * qhat * FFT3(rho[1] zero-padded on
 * the doubled grid))[:nx,:ny,:nz]; rho is the (2,nx,ny,nz) density, only
 * rho[1] is read.  work: (2nx)(2ny)(2nz) complex doubles. */
int hfb_poisson(hfb_ctx* ctx, const double* rho, double* wcoul, double e2, void* work, void* stream);

/* Largest doubled Coulomb grid (points) of hfb_poisson_noalloc: 96^3, i.e.
 * physical grids up to 48^3 (14.2 MB of device memory per GPU). */
#define HFB_POISSON_SCRATCH_MAX (96LL * 96 * 96)

/* hfb_poisson without a work argument and without any allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The doubled-grid work array is a static device array
 * of HFB_POISSON_SCRATCH_MAX complex doubles on each GPU, held until the
 * process ends; a context whose doubled grid is larger gets HFB_E_SIZE.
 * Every call on one GPU shares that array: calls that overlap in time on
 * one GPU (other streams, other threads, other contexts) race and must be
 * ordered by the caller.  Nothing in it needs to start at zero. */
int hfb_poisson_noalloc(hfb_ctx* ctx, const double* rho, double* wcoul, double e2, void* stream);

/* ---------------------------------------------------------------------
 * Gradient step (static.py:93-259: grstep, grstep_helper, e0dmp_gt_zero;
 * levels.py:397-437: laplace)
 * --------------------------------------------------------------------- */

/* Per-state weights of grstep_helper (static.py:174-185): w = wstates*wocc,
 * wuv = wstates*wguv*pairwg (w=1, wuv=0 when tbcs), and the preconditioner
 * parameters of e0dmp_gt_zero (static.py:131-143): pw = weight of the
 * (E0 - h2ma k^2) term, pshift = 1/2 max(wuv,.1) max_r v_pair[iq].
 * vpair_max (2 doubles, e.g. E + HFB_EN_VMAX as written by hfb_skyrme)
 * gives max_r v_pair[iq]; when it is null, v_pair (2, nvox) double is
 * scanned instead. */
int hfb_grstep_weights(int nstates, int npsi0, int tbcs, const double* wocc, const double* wstates,
                       const double* wguv, const double* pairwg, const double* v_pair, long long nvox,
                       const double* vpair_max, double* w, double* wuv, double* pw, double* pshift, void* stream);

/* Number of partial-sum tiles per field written by hfb_precond_step. */
int hfb_step_ntiles(hfb_ctx* ctx, int prec);

/* In place psi <- psi - x0dmp * laplace(r) for nstates two-spinors
 * (e0dmp_gt_zero, static.py:93-146, with laplace levels.py:397-437):
 *   h_exp = wxyz * hsum[2s+1]   (Re<psi|ps1>, from hfb_hpsi's partials)
 * hsum: per-state sums (S, 2) when hsum_ntile == 0, else hfb_hpsi's
 * partials (S, hsum_ntile, 2), added in tile order inside the kernels.
 *   r = ps1 - lagrange          if use_lagrange   (iteration > 1 and !tbcs)
 *   r = ps1 - h_exp psi         otherwise
 *   laplace: FFT3 r / (pw (e0dmp - h2ma k^2) + pshift), Nyquist kept.
 * fpart (optional, double): per state and tile (|r|^2, |lagrange - h_exp
 * psi|^2) partial sums for sp_efluct1/2 (static.py:120-129); reduce with
 * hfb_reduce_partials.  work: one spinor block of prec's complex type. */
int hfb_precond_step(hfb_ctx* ctx, int prec, void* psi, const void* ps1, const void* lagrange, int nstates,
                     const double* hsum, int hsum_ntile, double wxyz, int use_lagrange, const double* pw, const double* pshift,
                     double e0dmp, double h2ma, double x0dmp, double* fpart, void* work, void* stream);

/* Scalar tail of grstep (static.py:204-213, 250-257): sp_energy <- wxyz *
 * hsum[2s]; denerg = (spe_old - spe_new)/|spe_new|; sp_efluct1/2 <-
 * sqrt(fsum) when fsum is given (mprint iterations); acc[0] += sum
 * sp_efluct1 (static.sumflu), acc[1] += sum wocc*wstates*denerg
 * (static.delesum).  hsum as in hfb_precond_step; fsum: (S, 2) totals
 * (fsum_ntile == 0) or hfb_precond_step's fpart with fsum_ntile = twice its
 * tile count (both spin fields of a state), scaled by wxyz here. */
int hfb_grstep_finish(int nstates, const double* hsum, int hsum_ntile, double wxyz, double* sp_energy,
                      const double* wocc, const double* wstates, const double* fsum, int fsum_ntile,
                      double* sp_efluct1, double* sp_efluct2, double* acc, void* stream);

/* ---------------------------------------------------------------------
 * diagstep (static.py:261-577): Gram matrices on MFMA, small dense
 * eigenproblems, rotation of the state block
 * --------------------------------------------------------------------- */

/* Doubles of partial-sum workspace hfb_gram / hfb_gram2 need for one block
 * of nst states of K = 2*nx*ny*nz complex entries. */
int hfb_gram_workspace(int nst, long long K);

/* out[m] = scale * A^H B[m] (m < nm <= 4), A and B[m] (nst, K) blocks of
 * prec's complex type, out (nm, nst, nst) complex128: the psi_2d^H psi_2d,
 * psi_2d^H hampsi, psi_2d^H hmfpsi and psi_2d^H delpsi products of
 * loewdin_orthonormalize (static.py:470), diagstep (static.py:305-310) and
 * construct_hfb_matrices (static.py:503-529), with scale = wxyz. */
int hfb_gram(int prec, const void* A, const void* const* B, int nm, int nst, long long K, double scale,
             double* part, double* out, void* stream);

/* hfb_gram for both isospin blocks in one launch: A[iq], B[iq*4 + m],
 * nst[iq], part[iq], out[iq]; a block with nst[iq] == 0 is skipped. */
int hfb_gram2(int prec, const void* const* A, const void* const* B, int nm, const int* nst, long long K,
              double scale, double* const* part, double* const* out, void* stream);

/* hfb_gram2 with hfb_grstep_finish (job, may be null) in one extra
 * workgroup of the same launch: the gradient step's scalar tail
 * (static.py:204-213, 250-257) next to the Gram sums of diagstep, which
 * read none of what it writes. */
int hfb_gram2_ex(int prec, const void* const* A, const void* const* B, int nm, const int* nst, long long K,
                 double scale, double* const* part, double* const* out, const hfb_finish_job* job, void* stream);

/* Loewdin orthonormalisation, optional subspace diagonalisation and the HFB
 * matrices for both isospins (static.py:463-482 loewdin_orthonormalize with
 * its eigenvalue clamp; static.py:296-316 lower-triangle eigh of
 * lambda_lin; static.py:484-577 construct_hfb_matrices).  mats[iq*4 + m]:
 * m=0 S, 1 H_old, 2 lambda_lin (or null), 3 G_old (or null), each
 * nst_iq x nst_iq complex128 from hfb_gram2.  uo[iq*2] receives the
 * unitary U (nst x nst complex128, row-major: psi_new = psi_old U, the
 * input of hfb_rotate), uo[iq*2+1] M2 = U lambda+, which maps the states
 * before the rotation to lagrange.  Writes sp_norm,
 * sp_energy, deltaf (ipair), hmatrix, gapmatrix, symcond, lambda_save
 * (2, nmax, nmax) complex128 and efluct1q/2q (2), efluct1/2 (1).
 * nst0, nst1 <= 256 (HFB_E_SIZE above).  work: null when both blocks
 * have at most 36 states (the work matrices then live in LDS), else
 * 10 * nmax^2 complex doubles.  stats (optional, 4
 * doubles): Jacobi sweeps of the lambda_lin eigensolve per isospin, then
 * Newton-Schulz iterations per isospin (0: the overlap was within 1e-3 of
 * the identity in Frobenius norm and the binomial series to fourth order
 * gave S^{-1/2}; -1: fell back to the eigensolve).  jacobi_tol: the
 * relative off-diagonal Frobenius norm at which the lambda_lin eigensolve
 * stops (0: 1e-13, the accuracy of FP64 subspace matrices; the mixed path
 * passes 1e-8 for matrices of FP32 states, whose entries carry about 1e-7;
 * HFB_E_ARG outside [0, 1)).
 * efluct1/2 are combined over the isospins inside the same launch (an
 * arrival counter of the context).  Above 36 states without tdiag the
 * series, the products and the efluct combination run as a sequence of
 * launches over many compute units (per-isospin control words of the
 * context choose the blocks that stay on the one-workgroup kernel), so
 * calls on one context must not overlap in time. */
int hfb_diag_small(hfb_ctx* c, int nst0, int nst1, int nmax, const double* const* mats, int tdiag, int ipair, int tbcs,
                   int construct, const double* wocc, const double* wstates, const double* wguv,
                   const double* pairwg, double* sp_norm, double* sp_energy, double* deltaf, double* hmatrix,
                   double* gapmatrix, double* symcond, double* lambda_save, double* efluct1q, double* efluct2q,
                   double* efluct1, double* efluct2, double* const* uo, double* work, double* stats,
                   double jacobi_tol, void* stream);

/* loewdin_orthonormalize (static.py:463-482) of one block from its overlap
 * S = wxyz psi_2d^H psi_2d (nst x nst complex128, e.g. from hfb_gram): U
 * (nst x nst complex128, row-major, psi_new = psi_2d U) = S^{-1/2} and
 * sp_norm = Re diag S.  Same solver as hfb_diag_small (binomial series or
 * Newton-Schulz near the identity, else the eigen-decomposition with the
 * reference's eigenvalue clamp).  nst <= 256; work: null up to 36 states,
 * else 5 nst^2 complex doubles. */
int hfb_loewdin(int nst, const double* S, double* U, double* sp_norm, double* work, void* stream);

/* construct_hfb_matrices (static.py:484-577) of isospin iq (states start ..
 * start+nst-1, already rotated by U) with the caller's unitary U and
 * H_old = wxyz psi_old^H hmfpsi, G_old = wxyz psi_old^H delpsi (ipair;
 * may be null otherwise), all nst x nst complex128: writes rows iq of
 * hmatrix, gapmatrix (diagonal only when tbcs), symcond, lambda_save
 * (2, nmax, nmax), sp_energy and deltaf (ipair) of those states,
 * efluct1q/2q[iq], efluct1 = max and efluct2 = mean over both isospins'
 * efluct1q/2q, and lam = lambda+ (nst x nst) so that lagrange = psi lam
 * (e.g. hfb_rotate).  nst <= 256; work as for hfb_loewdin. */
int hfb_hfb_matrices(hfb_ctx* c, int iq, int start, int nst, int nmax, const double* U, const double* Hold,
                     const double* Gold, int ipair, int tbcs, const double* wocc, const double* wstates,
                     const double* wguv, const double* pairwg, double* sp_energy, double* deltaf, double* hmatrix,
                     double* gapmatrix, double* symcond, double* lambda_save, double* efluct1q, double* efluct2q,
                     double* efluct1, double* efluct2, double* lam, double* work, void* stream);

/* hfb_natural (static.py:920-1070): gradient sub-iterations of the
 * canonical basis of isospin iq (states start .. start+nst-1; the
 * reference slices by nneut/nprot) in configuration space.  phase 0 runs
 * the iternat sub-iterations: cm <- QR(cm - dampstep * symcond) (Householder
 * with LAPACK's conventions, the Q of jnp.linalg.qr), htran = cm^H h cm,
 * gaptran likewise, sp_energy and deltaf of the block, the one-isospin
 * pairing update between sub-iterations (static.py:1170-1336; its brentq
 * call always raises, so the Fermi energy is the midpoint of levels N-1 and
 * N), symcond <- w hcm - wuv gapcm - cm lambda.  The caller then runs the
 * full pair (hfb_bcs) of the last sub-iteration when ipair, and phase 1
 * writes lambda_save, hmatrix, gapmatrix of iq, cm_out = cm and m2_out =
 * cm htran (nst x nst complex128) for psi <- psi cm and lagrange <- psi cm
 * htran (hfb_rotate).  hmatrix..lambda_save: (2, nmax, nmax) complex128;
 * eferm..avg: (2) doubles.  work: 10 nst^2 + nst complex doubles, kept
 * between the phases.  nst <= 256. */
int hfb_natural(int phase, int iq, int start, int nst, int nmax, int iternat, int ipair, double delstepbas,
                double e0bas, double npart, double pair_cut, double state_cut, double softcut, double* hmatrix,
                double* gapmatrix, double* symcond, double* lambda_save, double* sp_energy, double* deltaf,
                double* wocc, double* wguv, double* pairwg, double* wstates, double* eferm, double* epair,
                double* avdelt, double* avdeltv2, double* avg, double* work, double* cm_out, double* m2_out,
                void* stream);

/* lastdiag (static.py:1072-1168) of isospin iq over states start ..
 * start+nst-1: eigenvalues of the Hermitian part of (h - eferm) xi + Delta
 * eta in the canonical basis (xi_ij = -sqrt(v_i v_j) + sqrt((1-v_i)(1-v_j)),
 * eta_ij = sqrt(v_i(1-v_j)) + sqrt(v_j(1-v_i)), v = wocc) in ascending
 * order into qp_energy, the norms of their lower components into qp_norm
 * (nst doubles each).  work: nst (nst + 1) / 2 + 2 nst^2 complex doubles. */
int hfb_lastdiag(int iq, int start, int nst, int nmax, double eferm, const double* hmatrix, const double* gapmatrix,
                 const double* wocc, double* qp_energy, double* qp_norm, double* work, void* stream);

/* out1 = psi_2d U1 and out2 = psi_2d U2 over the nst states of one block
 * (static.py:316 psi_2d @ unitary; static.py:557 lagrange = psi_new @
 * lambda+), psi (nst, K), U1/U2 (nst, nst) complex128.  out1 may alias psi
 * (in place); U2/out2 may be null.  A workgroup stages 64 elements of
 * every state in LDS: HFB_E_SIZE above 300 states (FP32) or 150 (FP64). */
int hfb_rotate(int prec, const void* psi, const double* U1, const double* U2, void* out1, void* out2, int nst,
               long long K, void* stream);

/* hfb_rotate for both isospin blocks in one launch (the two rotations of
 * static.py:316 and static.py:557 for neutrons and protons): arrays of two
 * entries indexed by isospin; a block with nst[iq] == 0 is skipped. */
int hfb_rotate2(int prec, const void* const* psi, const double* const* U1, const double* const* U2,
                void* const* out1, void* const* out2, const int* nst, long long K, void* stream);

/* ---------------------------------------------------------------------
 * Densities and single-particle properties (densities.py:177-337,
 * inout.py:64-149)
 * --------------------------------------------------------------------- */

/* Sizes in doubles of the dout scratch of hfb_density and of its per-state
 * partial sums.  The layout of dout depends on prec and the grid (FP32
 * states on 32^3 grids fill half of it with FP32 partials); the finish
 * calls take the same prec. */
int hfb_density_sizes(hfb_ctx* ctx, int prec, long long* dout_doubles, long long* spart_doubles_per_state);

/* One sweep over the states of both isospins (nst0 + nst1 of them) that
 * accumulates, with weights w = wocc*wstates and wuv = wguv*pairwg*wstates,
 * rho, chi, tau, current, sdens and sodens (add_density with
 * compute_single_contribution, densities.py:177-337) into dout when
 * want_dens, and the per-state sums of sp_properties (inout.py:64-149:
 * kinetic, orbital about cmtot[3], spin, parity) into spart when want_sp. */
int hfb_density(hfb_ctx* ctx, int prec, const void* psi, int nst0, int nst1, const double* wocc,
                const double* wstates, const double* wguv, const double* pairwg, const double* cmtot,
                int want_dens, int want_sp, double* dout, double* spart, void* stream);

/* dout (written by hfb_density with the same prec) -> the Densities fields
 * rho, chi, tau (2,n3), current, sdens, sodens (2,3,n3).  mix: x <-
 * addnew*x_new + (1-addnew)*x for rho, tau, chi (static.py:1841-1848); the
 * other fields are replaced. */
int hfb_density_finish(hfb_ctx* ctx, int prec, const double* dout, double* rho, double* chi, double* tau,
                       double* current, double* sdens, double* sodens, int mix, double addnew, void* stream);

/* hfb_density_finish and hfb_sp_finish in one launch (the loop's density
 * update and single-particle properties, static.py:1800-1823 with
 * inout.py:64-149); arguments as there. */
int hfb_density_sp_finish(hfb_ctx* ctx, int prec, const double* dout, double* rho, double* chi, double* tau,
                          double* current, double* sdens, double* sodens, int mix, double addnew,
                          const double* spart, int nst0, int nst1, double h2m0, double h2m1, double* sp_kinetic,
                          double* sp_orbital, double* sp_spin, double* sp_parity, void* stream);

/* spart -> sp_kinetic (S), sp_orbital (S,3), sp_spin (S,3), sp_parity (S)
 * with the wxyz and h2m_q factors of inout.py:140-148. */
int hfb_sp_finish(hfb_ctx* ctx, const double* spart, int nst0, int nst1, double h2m0, double h2m1,
                  double* sp_kinetic, double* sp_orbital, double* sp_spin, double* sp_parity, void* stream);

/* ---------------------------------------------------------------------
 * Potentials and energies (meanfield.py:163-391, energies.py:90-371)
 * --------------------------------------------------------------------- */

/* Doubles of the workspace of hfb_density_derivs and hfb_meanfield_derivs
 * (the terms of the multi-term fields, summed in term order afterwards). */
long long hfb_deriv_workspace(hfb_ctx* ctx);

/* Derived density fields shared by skyrme and integ_energy, each per
 * isospin: divj = div sodens (meanfield.py:195-200), lap = Laplacian rho
 * (meanfield.py:227-231), grad = grad rho (meanfield.py:262-268), curls =
 * curl sdens (meanfield.py:279-289), curlj = curl current
 * (meanfield.py:304-314), with the sder/sder2 matrices of grids.py:36-63.
 * work: hfb_deriv_workspace doubles. */
int hfb_density_derivs(hfb_ctx* ctx, const double* rho, const double* sdens, const double* current,
                       const double* sodens, double* divj, double* lap, double* grad, double* curls,
                       double* curlj, double* work, void* stream);

/* hfb_density_derivs and hfb_poisson of rho (coulomb.py:54-75) in one call,
 * for the SCF loop: on cubic 32^3 and 48^3 grids the sums of the
 * multi-term derivative fields ride in the Coulomb x-pencil launch (four
 * launches instead of five); elsewhere the two entry points in turn.  work
 * as for hfb_density_derivs, pwork as for hfb_poisson. */
int hfb_density_derivs_coulomb(hfb_ctx* ctx, const double* rho, const double* sdens, const double* current,
                               const double* sodens, double* divj, double* lap, double* grad, double* curls,
                               double* curlj, double* work, double e2, double* wcoul, void* pwork, void* stream);

/* divaq = div aq and dbmass = grad bmass (meanfield.py:329-344); work as
 * for hfb_density_derivs. */
int hfb_meanfield_derivs(hfb_ctx* ctx, const double* aq, const double* bmass, double* divaq, double* dbmass,
                         double* work, void* stream);

/* Workgroups of the skyrme and energy kernels: hfb_skyrme's epart scratch
 * holds 3x, hfb_integ_energy's 20x that many doubles. */
int hfb_skyrme_nblocks(hfb_ctx* ctx);

/* Mean-field potentials upot, bmass, v_pair (2,n3), wlspot, aq, spot
 * (2,3,n3) from the densities, the derived fields and wcoul (skyrme,
 * meanfield.py:163-391, incl. the DDDI rearrangement and ecorrp into
 * E[HFB_EN_ECORRP]).  The maximum of v_pair per isospin goes to
 * E[HFB_EN_VMAX..+1] (for hfb_grstep_weights).  The *32 outputs are
 * float copies for the FP32 spinor kernels and may be null.  finish == 0
 * leaves ecorrp and the v_pair maxima as partials in epart for
 * hfb_energy_tail. */
int hfb_skyrme(hfb_ctx* ctx, const hfb_force* f, const double* rho, const double* tau, const double* chi,
               const double* current, const double* sdens, const double* divj, const double* lap,
               const double* grad, const double* curls, const double* curlj, const double* wcoul, double* upot,
               double* bmass, double* wlspot, double* aq, double* spot, double* v_pair, float* upot32,
               float* bmass32, float* wlspot32, float* aq32, float* spot32, float* v_pair32, double* epart,
               double* E, int finish, void* stream);

/* Energy-functional integrals of integ_energy (energies.py:90-306) into the
 * slots of E, incl. the C-coupling split and ehfint; pairing energies are
 * read from E[HFB_EN_EPAIR..+1].  finish == 0 leaves the block partials
 * in epart for hfb_energy_tail. */
int hfb_integ_energy(hfb_ctx* ctx, const hfb_force* f, const double* rho, const double* tau,
                     const double* current, const double* sdens, const double* divj, const double* lap,
                     const double* curlj, const double* wcoul, double* epart, double* E, int finish,
                     void* stream);

/* hfb_skyrme and hfb_integ_energy with finish == 0 in one launch (the
 * loop body's skyrme and integ_energy, meanfield.py:163-391 and
 * energies.py:90-306): the potentials, and the partials of ecorrp, the
 * v_pair maxima (sky_part, 3x hfb_skyrme_nblocks doubles) and of the
 * integrals (en_part, 20x) for hfb_energy_tail. */
int hfb_skyrme_energy(hfb_ctx* ctx, const hfb_force* f, const double* rho, const double* tau, const double* chi,
                      const double* current, const double* sdens, const double* divj, const double* lap,
                      const double* grad, const double* curls, const double* curlj, const double* wcoul,
                      double* upot, double* bmass, double* wlspot, double* aq, double* spot, double* v_pair,
                      float* upot32, float* bmass32, float* wlspot32, float* aq32, float* spot32, float* v_pair32,
                      double* sky_part, double* en_part, void* stream);

/* The scalar tail of the SCF loop body (static.py:1825-1833) in one
 * launch: ecorrp and the v_pair maxima from hfb_skyrme's partials
 * (sky_part, finish == 0), the integrals of integ_energy from
 * hfb_integ_energy's partials (en_part, finish == 0; current as given
 * there), then sum_energy (energies.py:309-371) on the single-particle
 * properties.  No single reference function corresponds: it finishes
 * skyrme, integ_energy and sum_energy together. */
int hfb_energy_tail(hfb_ctx* ctx, const hfb_force* f, const double* sky_part, const double* en_part,
                    const double* current, int nstates, const double* wocc, const double* wstates,
                    const double* sp_kinetic, const double* sp_energy, const double* sp_orbital,
                    const double* sp_spin, double* E, void* stream);

/* ehf, tke, orbital/spin/total angular momentum of sum_energy
 * (energies.py:309-371) into E. */
int hfb_sum_energy(int nstates, const double* wocc, const double* wstates, const double* sp_kinetic,
                   const double* sp_energy, const double* sp_orbital, const double* sp_spin, double* E,
                   void* stream);

/* ---------------------------------------------------------------------
 * Constraints on the quadrupole moments and the principal axes
 * (constraint.py:199-344).  Fields: (nc, nvox) doubles, isoscalar (one
 * field for both isospins), nc <= 8.
 * --------------------------------------------------------------------- */

/* upot (2, nvox) -= sum_i lambda_i field_i for both isospins
 * (add_constraint_potential, constraint.py:199-213, as called from skyrme,
 * meanfield.py:246-248); upot32 (optional): the float32 copy of upot,
 * refreshed from the result. */
int hfb_constraint_potential(int nc, long long nvox, const double* fields, const double* lambda, double* upot,
                             float* upot32, void* stream);

/* Expectation values <Q_i> = wxyz sum field_i (rho_n + rho_p), variances
 * |<Q_i^2> - <Q_i>^2 / N| and the multiplier update of constraint.py:216-289
 * in one launch: old[i] <- actual[i] (before_constraint_step), then
 * lambda[i] += -qeps max(e0act,1)/x0act (<Q_i> - old[i])/(2 var + d0),
 * qcorr[i] = c0 (<Q_i> - goal[i])/(2 var + d0), actual[i] = <Q_i>,
 * var[i] = var.  rho: (2, nvox). */
int hfb_constraint_tune(int nc, long long nvox, double wxyz, const double* fields, const double* rho, double qeps,
                        double d0, double c0, double e0act, double x0act, double* lambda, const double* goal,
                        double* actual, double* old, double* var, double* qcorr, void* stream);

/* psi_s (nstates, 2, nvox) *= exp(-wocc_s wstates_s sum_i qcorr_i
 * field_i) in place (apply_wavefunction_corrections, constraint.py:279-318). */
int hfb_constraint_correct(int prec, void* psi, int nstates, long long nvox, int nc, const double* fields,
                           const double* qcorr, const double* wocc, const double* wstates, void* stream);

/* ---------------------------------------------------------------------
 * BCS pairing (pairs.py:25-220)
 * --------------------------------------------------------------------- */

/* pair for both isospins: if const_gap >= 0 every deltaf is first set to
 * it (pairgap, pairs.py:135-137); then per isospin the Fermi energy of
 * sum clip(v2) wstates = N_q (the reference's brentq on [-100,100],
 * pairs.py:47-61; here a bisection to the same 1e-14 tolerance), v2, uv,
 * wocc, soft cutoffs, and eferm, epair, avdelt, avdeltv2, avg (2) into
 * their arrays and epair into E[HFB_EN_EPAIR..+1] (pairs.py:66-131).
 * status[0] is set to 1 when the bracket of either isospin has no sign
 * change (the reference's brentq ValueError); it is never cleared here.
 * Blocks of any size (levels staged in LDS up to 512 states). */
int hfb_bcs(int nst0, int nst1, double nneut, double nprot, const double* pair_cutoff,
            const double* state_cutoff, double softcut, double const_gap, const double* sp_energy,
            double* deltaf, double* wocc, double* wguv, double* pairwg, double* wstates, double* eferm,
            double* epair, double* avdelt, double* avdeltv2, double* avg, double* E, int* status, void* stream);

/* ---------------------------------------------------------------------
 * Energy-functional backward.  No counterpart in the reference: the
 * derivatives of ehfint (energies.py:90-306, summed at energies.py:267)
 * with respect to the densities and to the couplings t0..t3, x0..x3,
 * alpha (= power), W0 (= t4) of forces.py:60-143.
 * --------------------------------------------------------------------- */

/* Blocks of the per-point kernel; part needs 72 doubles per block. */
int hfb_energy_bwd_nblocks(hfb_ctx* ctx);

/* forces: nbatch (<= 8) host structs (the derived couplings of each set);
 * tx: device (nbatch, 10) raw couplings (t0, t1, t2, t3, x0, x1, x2, x3,
 * power, t4); gout: device (nbatch) upstream gradients dL/dE_b.  Inputs as
 * for hfb_integ_energy plus grad rho and curl s (hfb_density_derivs).
 * Outputs (each may be null): d_rho, d_tau (2,n3), d_sodens, d_current,
 * d_sdens (2,3,n3) = sum_b gout_b dE_b/d(field), which equal wxyz times
 * the mean fields U, B, W, (spin part of) A and S of meanfield.py:178-324;
 * d_coupling (device, nbatch x 10) = gout_b dE_b/d(coupling). */
int hfb_energy_bwd(hfb_ctx* ctx, const hfb_force* forces, const double* tx, const double* gout, int nbatch,
                   const double* rho, const double* tau, const double* sodens, const double* current,
                   const double* sdens, const double* divj, const double* lap, const double* grad,
                   const double* curls, const double* curlj, const double* wcoul, double* d_rho, double* d_tau,
                   double* d_sodens, double* d_current, double* d_sdens, double* d_coupling, double* part,
                   void* stream);

#ifdef __cplusplus
}
#endif
#endif
