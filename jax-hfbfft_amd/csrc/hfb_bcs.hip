// BCS occupations for both isospins (pairs.py:25-131), one workgroup each.
//
// The Fermi energy solves sum_k clip(v2_k(lambda), 1e-10, 1-1e-10) ws_k = N
// on [-100, 100] (pairs.py:47-61, brentq with xtol 1e-14).  The particle
// number is monotone in lambda, so the root is bracketed by a parallel
// bisection: each round evaluates the sum at one trial point per thread and
// keeps the sub-interval where the sign changes (the first trial point past
// it found by a wavefront ballot), shrinking the bracket by the thread
// count per round until it is below 1e-15 (brentq's own tolerance is
// 1e-14 + 4 eps |lambda|).  A bracket without a sign change sets
// status = 1 (brentq raises ValueError there).
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

// The part of sum_k clip(v2_k(lambda)) ws_k over the levels k = sub,
// sub + BCS_SPLIT, ...; v2 = (1 - edif / sqrt(edif^2 + d^2)) / 2 with the
// reference's 1e-20 floor on the square root, 1/sqrt by the refined hardware
// estimate (a few ulp).
constexpr int BCS_NT = 1024;  // threads: BCS_SPLIT per trial point
constexpr int BCS_SPLIT = 4;
__device__ __forceinline__ double rsq_refined(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  return y * fma(-0.5 * x * y, y, 1.5);
}
__device__ double particle_number_part(double lam, const double* e, const double* d, const double* ws, int n,
                                       int sub) {
  double s = 0.0;
#pragma unroll 4
  for (int k = sub; k < n; k += BCS_SPLIT) {
    const double edif = e[k] - lam;
    const double q2 = fmax(fma(edif, edif, d[k] * d[k]), 1.0e-40);
    double v2 = 0.5 - 0.5 * edif * rsq_refined(q2);
    v2 = fmin(fmax(v2, 1.0e-10), 1.0 - 1.0e-10);
    s = fma(v2, ws[k], s);
  }
  return s;
}
// the sum over the BCS_SPLIT consecutive lanes of a trial point
__device__ __forceinline__ double particle_number_lanes(double lam, const double* e, const double* d,
                                                        const double* ws, int n) {
  double s = particle_number_part(lam, e, d, ws, n, threadIdx.x & (BCS_SPLIT - 1));
  s += lane_xor<1>(s);
  s += lane_xor<2>(s);
  return s;
}

__global__ void __launch_bounds__(BCS_NT) bcs_kernel(BcsArgs a) {
  const int iq = blockIdx.x;
  const BcsIso b = a.iso[iq];
  const int n = b.nst;
  const int tid = threadIdx.x;
  // the levels of a block up to BCS_LDS states are staged in LDS, larger
  // blocks are read from global memory (L2) by every trial evaluation
  __shared__ double e_l[BCS_LDS], dl_l[BCS_LDS], ws_l[BCS_LDS];
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
  __syncthreads();
  __shared__ double fend[2];
  __shared__ int wfirst[BCS_NT / 64];
  // the bracket's end points (lanes 0-3 and 4-7)
  if (tid < 2 * BCS_SPLIT) {
    const double fe = particle_number_lanes(tid < BCS_SPLIT ? -100.0 : 100.0, e, dl, ws, n) - b.target;
    if ((tid & (BCS_SPLIT - 1)) == 0) fend[tid / BCS_SPLIT] = fe;
  }
  __syncthreads();
  if (tid == 0) {
    lo = -100.0;
    hi = 100.0;
    bad = 0;
    const double fa = fend[0], fb = fend[1];
    flo = fa;
    if (fa * fb > 0.0) bad = 1;
    if (fa == 0.0) hi = lo;
    if (fb == 0.0) lo = hi;
  }
  __syncthreads();
  constexpr int npt = BCS_NT / BCS_SPLIT;  // trial points per round
  const int pt = tid / BCS_SPLIT;
  for (int round = 0; round < 12 && !bad; ++round) {
    const double l0 = lo, h0 = hi;
    if (h0 - l0 <= 1.0e-15 * fmax(1.0, fabs(l0))) break;
    // trial points strictly inside (l0, h0), increasing with the index
    const double x = l0 + (h0 - l0) * double(pt + 1) / double(npt + 1);
    const double f = particle_number_lanes(x, e, dl, ws, n) - b.target;
    // the first trial point that is a root or lies past the sign change
    // (the particle number grows with lambda): a ballot per wavefront over
    // the first lane of each point, the smallest index over the wavefronts
    const bool past = f == 0.0 || ((f < 0.0) != (flo < 0.0));
    const unsigned long long m = __ballot(past) & 0x1111111111111111ull;
    if ((tid & 63) == 0) wfirst[tid >> 6] = m ? (tid >> 6) * (64 / BCS_SPLIT) + __builtin_ctzll(m) / BCS_SPLIT : npt;
    __syncthreads();
    int t = npt;
    for (int w = 0; w < BCS_NT / 64; ++w) t = min(t, wfirst[w]);
    // new bracket (the in-order scan's result): a root at t closes it
    // there; else lo = the trial point before t (l0 for t = 0), hi = the
    // point at t (h0 when every point lies before the sign change).  One
    // thread writes it; x_{t-1} is recomputed with the same expression.
    if (tid == (t < npt ? t : 0) * BCS_SPLIT) {
      if (t == npt) {
        lo = l0 + (h0 - l0) * double(npt) / double(npt + 1);
        hi = h0;
      } else if (f == 0.0) {
        lo = x;
        hi = x;
      } else {
        lo = t > 0 ? l0 + (h0 - l0) * double(t) / double(npt + 1) : l0;
        hi = x;
      }
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
  __shared__ double red[4 * (BCS_NT / 64)];
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
  hipLaunchKernelGGL(bcs_kernel, dim3(2), dim3(BCS_NT), 0, st, a);
  return launch_status();
}
