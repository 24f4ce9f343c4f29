// BCS occupations for both isospins (pairs.py:25-131), one workgroup each.
//
// The Fermi energy solves sum_k clip(v2_k(lambda), 1e-10, 1-1e-10) ws_k = N
// on [-100, 100] (pairs.py:47-61, brentq with xtol 1e-14).  The particle
// number is monotone in lambda, so the root is bracketed by a parallel
// bisection: each round evaluates the sum at one trial point per thread and
// keeps the sub-interval where the sign changes, shrinking the bracket by
// the thread count per round until it is below 1e-15 (brentq's own
// tolerance is 1e-14 + 4 eps |lambda|).  A bracket without a sign change
// sets status = 1 (brentq raises ValueError there).
#include <math.h>

#include "hfb_ctx.h"

using namespace hfb;

namespace {

constexpr int BCS_LDS = 512;  // largest block whose levels are staged in LDS

struct BcsIso {
  int start, nst;
  double target;      // particle number N or Z
  double pair_cut;    // forces.pair_cutoff[iq]
  double state_cut;   // forces.state_cutoff[iq]
};

struct BcsArgs {
  BcsIso iso[2];
  double softcut;
  const double* sp_energy;
  const double* deltaf;
  double* wocc;
  double* wguv;
  double* pairwg;
  double* wstates;
  double* eferm;    // (2)
  double* epair;    // (2)
  double* avdelt;   // (2)
  double* avdeltv2; // (2)
  double* avg;      // (2)
  double* E;        // energy vector: epair copy
  int* status;
};

__device__ double particle_number(double lam, const double* e, const double* d, const double* ws, int n) {
  double s = 0.0;
  for (int k = 0; k < n; ++k) {
    const double edif = e[k] - lam;
    const double eq = fmax(sqrt(edif * edif + d[k] * d[k]), 1.0e-20);
    double v2 = 0.5 * (1.0 - edif / eq);
    v2 = fmin(fmax(v2, 1.0e-10), 1.0 - 1.0e-10);
    s += v2 * ws[k];
  }
  return s;
}

__global__ void __launch_bounds__(256) bcs_kernel(BcsArgs a) {
  const int iq = blockIdx.x;
  const BcsIso b = a.iso[iq];
  const int n = b.nst;
  const int tid = threadIdx.x;
  // the levels of a block up to BCS_LDS states are staged in LDS, larger
  // blocks are read from global memory (L2) by every trial evaluation
  __shared__ double e_l[BCS_LDS], dl_l[BCS_LDS], ws_l[BCS_LDS];
  __shared__ double fval[256];
  __shared__ double lo, hi, flo;
  __shared__ int bad;
  const bool staged = n <= BCS_LDS;
  const double* e = staged ? e_l : a.sp_energy + b.start;
  const double* dl = staged ? dl_l : a.deltaf + b.start;
  const double* ws = staged ? ws_l : a.wstates + b.start;
  if (staged) {
    for (int k = tid; k < n; k += blockDim.x) {
      e_l[k] = a.sp_energy[b.start + k];
      dl_l[k] = a.deltaf[b.start + k];
      ws_l[k] = a.wstates[b.start + k];
    }
  }
  if (tid == 0) {
    lo = -100.0;
    hi = 100.0;
    bad = 0;
  }
  __syncthreads();
  if (tid == 0) {
    const double fa = particle_number(lo, e, dl, ws, n) - b.target;
    const double fb = particle_number(hi, e, dl, ws, n) - b.target;
    flo = fa;
    if (fa * fb > 0.0) bad = 1;
    if (fa == 0.0) hi = lo;
    if (fb == 0.0) lo = hi;
  }
  __syncthreads();
  const int nt = blockDim.x;
  for (int round = 0; round < 12 && !bad; ++round) {
    const double l0 = lo, h0 = hi;
    if (h0 - l0 <= 1.0e-15 * fmax(1.0, fabs(l0))) break;
    // trial points strictly inside (l0, h0)
    const double x = l0 + (h0 - l0) * double(tid + 1) / double(nt + 1);
    fval[tid] = particle_number(x, e, dl, ws, n) - b.target;
    __syncthreads();
    if (tid == 0) {
      double nl = l0, nh = h0;
      const bool neg_lo = flo < 0.0;
      for (int t = 0; t < nt; ++t) {
        const double xt = l0 + (h0 - l0) * double(t + 1) / double(nt + 1);
        if (fval[t] == 0.0) {
          nl = nh = xt;
          break;
        }
        if ((fval[t] < 0.0) == neg_lo) {
          nl = xt;
        } else {
          nh = xt;
          break;
        }
      }
      lo = nl;
      hi = nh;
    }
    __syncthreads();
  }
  __shared__ double ef;
  if (tid == 0) {
    ef = 0.5 * (lo + hi);
    if (bad) a.status[0] = 1;  // either isospin
  }
  __syncthreads();
  const double lam = ef;
  // occupations (pairs.py:102-111) and soft cut-offs (pairs.py:88-99)
  double v[4] = {0.0, 0.0, 0.0, 0.0};  // sumuv, sumduv, sumv2, sumdv2
  for (int k = tid; k < n; k += blockDim.x) {
    const int s = b.start + k;
    double wsc = ws[k];
    if (b.pair_cut > 0.0)
      a.pairwg[s] = 1.0 / (1.0 + exp((e[k] - (lam + b.pair_cut)) / (a.softcut * b.pair_cut)));
    if (b.state_cut > 0.0) {
      wsc = 1.0 / (1.0 + exp((e[k] - (lam + b.state_cut)) / (a.softcut * b.state_cut)));
      a.wstates[s] = wsc;
    }
    const double edif = e[k] - lam;
    const double eq = fmax(sqrt(edif * edif + dl[k] * dl[k]), 1.0e-20);
    const double v2 = 0.5 * (1.0 - edif / eq);
    const double uv = sqrt(fmax(v2 - v2 * v2, 1.0e-6));
    const double wo = fmin(fmax(v2, 1.0e-10), 1.0 - 1.0e-10);
    a.wocc[s] = wo;
    a.wguv[s] = uv;
    const double vol = 0.5 * uv * wsc;
    v[0] += vol;
    v[1] += vol * dl[k];
    v[2] += wo * wsc;
    v[3] += dl[k] * wo * wsc;
  }
  __shared__ double red[4 * 4];
  block_sum<4>(v, red);
  if (tid == 0) {
    const double suv = fmax(v[0], 1.0e-20), sv2 = fmax(v[2], 1.0e-20);
    a.eferm[iq] = lam;
    a.epair[iq] = v[1];
    a.avdelt[iq] = v[1] / suv;
    a.avdeltv2[iq] = v[3] / sv2;
    a.avg[iq] = v[1] / (suv * suv);
    if (a.E) a.E[HFB_EN_EPAIR + iq] = v[1];
  }
}

__global__ void const_gap_kernel(double* deltaf, int ns, double val) {
  for (int s = threadIdx.x + blockIdx.x * blockDim.x; s < ns; s += blockDim.x * gridDim.x) deltaf[s] = val;
}

}  // namespace

// deltaf (if const_gap >= 0) is first set to const_gap for every state
// (pairs.py:135-137); then the BCS solve of pairs.py:66-131 per isospin.
extern "C" int hfb_bcs(int nst0, int nst1, double nneut, double nprot, const double* pair_cutoff,
                       const double* state_cutoff, double softcut, double const_gap, const double* sp_energy,
                       double* deltaf, double* wocc, double* wguv, double* pairwg, double* wstates, double* eferm,
                       double* epair, double* avdelt, double* avdeltv2, double* avg, double* E, int* status,
                       void* stream) {
  if (nst0 < 0 || nst1 < 0) return HFB_E_SIZE;
  if (!pair_cutoff || !state_cutoff || !sp_energy || !deltaf || !wocc || !wguv || !pairwg || !wstates || !eferm ||
      !epair || !avdelt || !avdeltv2 || !avg || !status)
    return HFB_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (const_gap >= 0.0) {
    hipLaunchKernelGGL(const_gap_kernel, dim3(1), dim3(256), 0, st, deltaf, nst0 + nst1, const_gap);
    const int rc = launch_status();
    if (rc != HFB_OK) return rc;
  }
  BcsArgs a{};
  a.iso[0] = {0, nst0, nneut, pair_cutoff[0], state_cutoff[0]};
  a.iso[1] = {nst0, nst1, nprot, pair_cutoff[1], state_cutoff[1]};
  a.softcut = softcut;
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
  a.E = E;
  a.status = status;
  hipLaunchKernelGGL(bcs_kernel, dim3(2), dim3(256), 0, st, a);
  return launch_status();
}
