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

/* Slots of the device energy vector (double[HFB_EN_LEN]), energies.py:6-43. */
enum {
  HFB_EN_EHFT = 0, HFB_EN_EHF0, HFB_EN_EHF1, HFB_EN_EHF2, HFB_EN_EHF3, HFB_EN_EHFLS, HFB_EN_EHFLSODD,
  HFB_EN_EHFC, HFB_EN_ECORC, HFB_EN_EHFINT, HFB_EN_E3CORR, HFB_EN_EZPE, HFB_EN_CRHO0, HFB_EN_CRHO1,
  HFB_EN_CDRHO0, HFB_EN_CDRHO1, HFB_EN_CTAU0, HFB_EN_CTAU1, HFB_EN_CDJ0, HFB_EN_CDJ1, HFB_EN_CJ0,
  HFB_EN_CJ1, HFB_EN_EHF, HFB_EN_TKE, HFB_EN_ORB /* 3 */, HFB_EN_SPIN = HFB_EN_ORB + 3,
  HFB_EN_JTOT = HFB_EN_SPIN + 3, HFB_EN_EFL1 = HFB_EN_JTOT + 3, HFB_EN_EFL2, HFB_EN_EFL1Q /* 2 */,
  HFB_EN_EFL2Q = HFB_EN_EFL1Q + 2, HFB_EN_ECORRP = HFB_EN_EFL2Q + 2, HFB_EN_SUMFLU, HFB_EN_DELESUM,
  HFB_EN_EPAIR /* 2 */, HFB_EN_LEN = 48
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

/* Number of partial-sum tiles per state written by hfb_hpsi. */
int hfb_hpsi_ntiles(hfb_ctx* ctx, int prec);

/* out[s*nv + j] = scale * sum_t partial[(s*ntile + t)*nv + j]. */
int hfb_reduce_partials(const double* partial, int nstates, int ntile, int nv, double scale,
                        double* out, void* stream);

/* Self-test of the FP64 MFMA 16x16x4 lane layout used by the Gram kernels:
 * a (16x4) and b (4x16) row-major doubles in, d = a*b (16x16) out. */
int hfb_selftest_mfma_f64(const double* a, const double* b, double* d, void* stream);

#ifdef __cplusplus
}
#endif
#endif
