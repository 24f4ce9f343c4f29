// Constraints on the quadrupole moments and the principal axes
// (constraint.py:199-344): the constraint potential subtracted from U_q,
// the expectation values with their variances and the multiplier update,
// and the exponential correction of the wave functions.  The constraint
// fields are isoscalar (the same field for both isospins), stored once:
// (nc, nvox) doubles.
#include <math.h>

#include "hfb_ctx.h"

using namespace hfb;

namespace {

constexpr int CT = 256;
constexpr int NC_MAX = 8;  // Q20, Q22 and six principal-axis constraints

// upot[iq][r] -= sum_i lambda_i f_i(r) for both isospins; the float32 copy
// of upot (mixed precision) is refreshed from the result.
__global__ void __launch_bounds__(CT) constraint_potential_kernel(int nc, int64_t nvox, const double* __restrict__ f,
                                                                  const double* __restrict__ lam, double* upot,
                                                                  float* upot32) {
  double l[NC_MAX];
#pragma unroll
  for (int i = 0; i < NC_MAX; ++i) l[i] = i < nc ? lam[i] : 0.0;
  for (int64_t r = (int64_t)blockIdx.x * CT + threadIdx.x; r < nvox; r += (int64_t)gridDim.x * CT) {
    double c = 0.0;
    for (int i = 0; i < nc; ++i) c = fma(l[i], f[i * nvox + r], c);
    for (int iq = 0; iq < 2; ++iq) {
      const double u = upot[iq * nvox + r] - c;
      upot[iq * nvox + r] = u;
      if (upot32) upot32[iq * nvox + r] = (float)u;
    }
  }
}

struct TuneArgs {
  int nc;
  int64_t nvox;
  double wxyz, qeps, d0, c0, e0act, x0act;
  const double* f;
  const double* rho;  // (2, nvox)
  double* lambda;
  const double* goal;
  double* actual;
  double* old;
  double* var;
  double* qcorr;
};

// Workgroup i: <Q_i> = wxyz sum f_i rho, <Q_i^2>, N = wxyz sum rho; the
// previous <Q_i> becomes old_i (the reference's before_constraint_step),
// then lambda_i and qcorr_i are updated (constraint.py:216-289).
__global__ void __launch_bounds__(CT) constraint_tune_kernel(TuneArgs a) {
  const int i = blockIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  const double* fi = a.f + i * a.nvox;
  for (int64_t r = threadIdx.x; r < a.nvox; r += CT) {
    const double rho = a.rho[r] + a.rho[a.nvox + r];
    const double q = fi[r];
    v[0] += q * rho;
    v[1] += q * q * rho;
    v[2] += rho;
  }
  __shared__ double red[3 * (CT / 64)];
  block_sum<3>(v, red);
  if (threadIdx.x == 0) {
    const double ex = a.wxyz * v[0], m2 = a.wxyz * v[1], numb = a.wxyz * v[2];
    const double var = fabs(m2 - ex * ex / numb);
    const double old = a.actual[i];
    a.old[i] = old;
    a.lambda[i] += -a.qeps * (fmax(a.e0act, 1.0) / a.x0act) * (ex - old) / (2.0 * var + a.d0);
    a.qcorr[i] = a.c0 * (ex - a.goal[i]) / (2.0 * var + a.d0);
    a.actual[i] = ex;
    a.var[i] = var;
  }
}

// psi_s(r) *= exp(-wocc_s wstates_s sum_i qcorr_i f_i(r)), both spin
// components (constraint.py:292-318).
template <typename T>
__global__ void __launch_bounds__(CT) constraint_correct_kernel(cx<T>* psi, int nstates, int64_t nvox, int nc,
                                                                const double* __restrict__ f,
                                                                const double* __restrict__ qcorr,
                                                                const double* __restrict__ wocc,
                                                                const double* __restrict__ wstates) {
  double q[NC_MAX];
#pragma unroll
  for (int i = 0; i < NC_MAX; ++i) q[i] = i < nc ? qcorr[i] : 0.0;
  const int s = blockIdx.y;
  const double w = wocc[s] * wstates[s];
  cx<T>* p = psi + (int64_t)s * 2 * nvox;
  for (int64_t r = (int64_t)blockIdx.x * CT + threadIdx.x; r < nvox; r += (int64_t)gridDim.x * CT) {
    double c = 0.0;
    for (int i = 0; i < nc; ++i) c = fma(q[i], f[i * nvox + r], c);
    const T e = (T)exp(-w * c);
    p[r] = cscale(p[r], e);
    p[nvox + r] = cscale(p[nvox + r], e);
  }
}

int grid_for(int64_t nvox) {
  const int64_t b = (nvox + CT - 1) / CT;
  return (int)(b < 1024 ? b : 1024);
}

}  // namespace

extern "C" int hfb_constraint_potential(int nc, long long nvox, const double* fields, const double* lambda,
                                        double* upot, float* upot32, void* stream) {
  if (nc <= 0 || nc > NC_MAX || nvox <= 0) return HFB_E_SIZE;
  if (!fields || !lambda || !upot) return HFB_E_ARG;
  hipLaunchKernelGGL(constraint_potential_kernel, dim3(grid_for(nvox)), dim3(CT), 0, (hipStream_t)stream, nc,
                     (int64_t)nvox, fields, lambda, upot, upot32);
  return launch_status();
}

extern "C" int hfb_constraint_tune(int nc, long long nvox, double wxyz, const double* fields, const double* rho,
                                   double qeps, double d0, double c0, double e0act, double x0act, double* lambda,
                                   const double* goal, double* actual, double* old, double* var, double* qcorr,
                                   void* stream) {
  if (nc <= 0 || nc > NC_MAX || nvox <= 0) return HFB_E_SIZE;
  if (!fields || !rho || !lambda || !goal || !actual || !old || !var || !qcorr || !(x0act != 0.0)) return HFB_E_ARG;
  TuneArgs a{nc, (int64_t)nvox, wxyz, qeps, d0, c0, e0act, x0act, fields, rho, lambda, goal, actual, old, var, qcorr};
  hipLaunchKernelGGL(constraint_tune_kernel, dim3(nc), dim3(CT), 0, (hipStream_t)stream, a);
  return launch_status();
}

extern "C" int hfb_constraint_correct(int prec, void* psi, int nstates, long long nvox, int nc, const double* fields,
                                      const double* qcorr, const double* wocc, const double* wstates, void* stream) {
  if (nc <= 0 || nc > NC_MAX || nvox <= 0 || nstates <= 0) return HFB_E_SIZE;
  if (!psi || !fields || !qcorr || !wocc || !wstates) return HFB_E_ARG;
  const dim3 grid(grid_for(nvox), nstates);
  if (prec == HFB_F64)
    hipLaunchKernelGGL(constraint_correct_kernel<double>, grid, dim3(CT), 0, (hipStream_t)stream, (cx<double>*)psi,
                       nstates, (int64_t)nvox, nc, fields, qcorr, wocc, wstates);
  else if (prec == HFB_F32)
    hipLaunchKernelGGL(constraint_correct_kernel<float>, grid, dim3(CT), 0, (hipStream_t)stream, (cx<float>*)psi,
                       nstates, (int64_t)nvox, nc, fields, qcorr, wocc, wstates);
  else
    return HFB_E_ARG;
  return launch_status();
}
