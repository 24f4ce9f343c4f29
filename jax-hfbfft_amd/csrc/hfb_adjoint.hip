// Backward of the energy functional ehfint (energies.py:90-306, summed at
// energies.py:267) with respect to the densities and the Skyrme couplings
// t0..t3, x0..x3, alpha (= power) and W0 (= t4).  The reference has no such
// routine; the derivatives follow from the functional term by term:
//
//   dE/d rho_q = wxyz [ b0 rho - b0' rho_q + b1 tau - b1' tau_q - b2 lap rho + b2' lap rho_q
//                       + rho^a ((b3 (a+2)/3 - 2 b3'/3) rho_q + b3 (a+2)/3 rho_q'
//                                - (a b3'/3) (rho_n^2 + rho_p^2) / rho)
//                       - (b4 + b4') div J_q - b4 div J_q'
//                       + [q = p] (W_coul - slate rho_p^(1/3)) ]
//   dE/d tau_q = wxyz (h2m_q + b1 rho - b1' rho_q)
//   dE/d J_q   = wxyz ((b4 + b4') grad rho_q + b4 grad rho_q')   (der1 is antisymmetric)
//   dE/d s_q   = wxyz (-(b4 + b4') curl j_q - b4 curl j_q')
//   dE/d j_q   = wxyz (-(b4 + b4') curl s_q - b4 curl s_q')     (curl is symmetric)
//
// which are the mean fields U, B, W, S of meanfield.py:178-324 and the
// spin part of A_q (ehfint has no j^2 term).  The coupling derivatives are
// the functional's integrals against each b-coupling, chained through the
// t/x definitions of forces.py:95-122.  A batch of coupling sets shares the
// densities: the density gradients are summed over the batch with the
// upstream weights, the coupling gradients are returned per set.
#include <math.h>

#include "hfb_ctx.h"

using namespace hfb;

namespace {

constexpr int MAXB = 8;  // coupling sets per launch (kernel-argument size)
constexpr int NI = 11;    // integrals per set
enum { I_B0, I_B0P, I_B1, I_B1P, I_B2, I_B2P, I_B3, I_B3P, I_ALPHA, I_B4, I_B4P };

struct AdjArgs {
  hfb_force f[MAXB];
  const double* gout;  // upstream gradient of each set's energy (device, nb)
  int nb;
  int64_t nvox;
  double wxyz;
  const double *rho, *tau, *sodens, *cur, *sdens;       // densities
  const double *divj, *lap, *grad, *curls, *curlj;      // derived fields
  const double* wcoul;
  double *d_rho, *d_tau, *d_sodens, *d_cur, *d_sdens;   // may be null
  double* part;  // [block][nb][NI]
};

__global__ void __launch_bounds__(256) energy_bwd_kernel(AdjArgs a) {
  const int64_t n = a.nvox;
  __shared__ double red[4 * NI];
  // integrals that do not depend on the coupling set, and the three that do
  double ic[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // I_B0 .. I_B2P, I_B4, I_B4P
  double ib[MAXB][3];                        // I_B3, I_B3P, I_ALPHA
#pragma unroll
  for (int b = 0; b < MAXB; ++b) ib[b][0] = ib[b][1] = ib[b][2] = 0.0;
  for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
    const double rq[2] = {a.rho[v], a.rho[n + v]};
    const double tq[2] = {a.tau[v], a.tau[n + v]};
    const double lq[2] = {a.lap[v], a.lap[n + v]};
    const double dq[2] = {a.divj[v], a.divj[n + v]};
    const double rt = rq[0] + rq[1], tt = tq[0] + tq[1], lt = lq[0] + lq[1], dt = dq[0] + dq[1];
    const double r2 = rq[0] * rq[0] + rq[1] * rq[1];
    double sodd_tot = 0.0, sodd_q = 0.0;  // s.curl j summed over q, and per-q diagonal
    for (int c = 0; c < 3; ++c) {
      const double c0 = a.curlj[(0 * 3 + c) * n + v], c1 = a.curlj[(1 * 3 + c) * n + v];
      const double s0 = a.sdens[(0 * 3 + c) * n + v], s1 = a.sdens[(1 * 3 + c) * n + v];
      sodd_tot += (s0 + s1) * (c0 + c1);
      sodd_q += s0 * c0 + s1 * c1;
    }
    ic[0] += 0.5 * rt * rt;                               // b0
    ic[1] += -0.5 * r2;                                   // b0'
    ic[2] += rt * tt;                                     // b1
    ic[3] += -(rq[0] * tq[0] + rq[1] * tq[1]);            // b1'
    ic[4] += -0.5 * rt * lt;                              // b2
    ic[5] += 0.5 * (rq[0] * lq[0] + rq[1] * lq[1]);       // b2'
    ic[6] += -rt * dt - sodd_tot;                         // b4
    ic[7] += -(rq[0] * dq[0] + rq[1] * dq[1]) - sodd_q;   // b4'
    const double lnr = rt > 0.0 ? log(rt) : 0.0;
    double g_rho[2] = {0.0, 0.0}, g_tau[2] = {0.0, 0.0};
    double g_J[2][3] = {}, g_s[2][3] = {}, g_j[2][3] = {};
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {  // compile-time b keeps ib[b][.] in registers
      if (b < a.nb) {
        const hfb_force& f = a.f[b];
        const double go = a.gout[b];
        const double rpow = pow(rt, f.power);
        ib[b][0] += rpow * rt * rt / 3.0;
        ib[b][1] += -rpow * r2 / 3.0;
        ib[b][2] += rpow * lnr * (f.b3 * rt * rt - f.b3p * r2) / 3.0;
        for (int q = 0; q < 2; ++q) {
          const int c = 1 - q;
          double u = f.b0 * rt - f.b0p * rq[q] + f.b1 * tt - f.b1p * tq[q] - f.b2 * lt + f.b2p * lq[q];
          u += rpow * ((f.b3 * (f.power + 2.0) / 3.0 - 2.0 * f.b3p / 3.0) * rq[q] +
                       f.b3 * (f.power + 2.0) / 3.0 * rq[c] - (f.b3p * f.power / 3.0) * r2 / (rt + 1.0e-25));
          u += -(f.b4 + f.b4p) * dq[q] - f.b4 * dq[c];
          if (q == 1 && f.tcoul) {
            u += a.wcoul[v];
            if (f.ex != 0) u += -f.slate * pow(rq[1], 1.0 / 3.0);
          }
          g_rho[q] += go * u;
          g_tau[q] += go * (f.h2m[q] + f.b1 * rt - f.b1p * rq[q]);
          for (int k = 0; k < 3; ++k) {
            const int64_t iqk = (int64_t)(q * 3 + k) * n + v, ick = (int64_t)(c * 3 + k) * n + v;
            g_J[q][k] += go * ((f.b4 + f.b4p) * a.grad[iqk] + f.b4 * a.grad[ick]);
            g_s[q][k] += go * (-(f.b4 + f.b4p) * a.curlj[iqk] - f.b4 * a.curlj[ick]);
            g_j[q][k] += go * (-(f.b4 + f.b4p) * a.curls[iqk] - f.b4 * a.curls[ick]);
          }
        }
      }
    }
    const double w = a.wxyz;
    for (int q = 0; q < 2; ++q) {
      if (a.d_rho) a.d_rho[q * n + v] = w * g_rho[q];
      if (a.d_tau) a.d_tau[q * n + v] = w * g_tau[q];
      for (int k = 0; k < 3; ++k) {
        const int64_t iqk = (int64_t)(q * 3 + k) * n + v;
        if (a.d_sodens) a.d_sodens[iqk] = w * g_J[q][k];
        if (a.d_sdens) a.d_sdens[iqk] = w * g_s[q][k];
        if (a.d_cur) a.d_cur[iqk] = w * g_j[q][k];
      }
    }
  }
  block_sum<8>(ic, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 8; ++k) a.part[((int64_t)blockIdx.x * (MAXB + 1)) * 8 + k] = ic[k];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    if (b < a.nb) {
      block_sum<3>(ib[b], red);
      if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) a.part[((int64_t)blockIdx.x * (MAXB + 1) + 1 + b) * 8 + k] = ib[b][k];
    }
  }
}

// Sum the block partials in a fixed order and chain them to the t/x
// couplings: tx[b] = (t0, t1, t2, t3, x0, x1, x2, x3, power, t4).
__global__ void energy_bwd_finish_kernel(const double* __restrict__ part, int nblk, double wxyz,
                                         const double* __restrict__ tx, const double* __restrict__ gout,
                                         double* __restrict__ d_coupling) {
  const int b = blockIdx.x;
  __shared__ double red[4 * NI];
  double I[NI];
  for (int k = 0; k < NI; ++k) I[k] = 0.0;
  for (int blk = threadIdx.x; blk < nblk; blk += blockDim.x) {
    const double* pc = part + (int64_t)blk * (MAXB + 1) * 8;
    const double* pb = pc + (1 + b) * 8;
    I[I_B0] += pc[0];
    I[I_B0P] += pc[1];
    I[I_B1] += pc[2];
    I[I_B1P] += pc[3];
    I[I_B2] += pc[4];
    I[I_B2P] += pc[5];
    I[I_B4] += pc[6];
    I[I_B4P] += pc[7];
    I[I_B3] += pb[0];
    I[I_B3P] += pb[1];
    I[I_ALPHA] += pb[2];
  }
  block_sum<NI>(I, red);
  if (threadIdx.x == 0) {
    const double* p = tx + 10 * b;
    const double t0 = p[0], t1 = p[1], t2 = p[2], t3 = p[3], x0 = p[4], x1 = p[5], x2 = p[6], x3 = p[7];
    const double s = wxyz * gout[b];
    double* d = d_coupling + 10 * b;
    // b0 = t0 (1 + x0/2), b0' = t0 (1/2 + x0)
    d[0] = s * (I[I_B0] * (1.0 + 0.5 * x0) + I[I_B0P] * (0.5 + x0));
    d[4] = s * (I[I_B0] * 0.5 * t0 + I[I_B0P] * t0);
    // b1 = (t1 (1 + x1/2) + t2 (1 + x2/2)) / 4,  b1' = (t1 (1/2 + x1) - t2 (1/2 + x2)) / 4
    // b2 = (3 t1 (1 + x1/2) - t2 (1 + x2/2)) / 8, b2' = (3 t1 (1/2 + x1) + t2 (1/2 + x2)) / 8
    d[1] = s * (I[I_B1] * (1.0 + 0.5 * x1) / 4.0 + I[I_B1P] * (0.5 + x1) / 4.0 + I[I_B2] * 3.0 * (1.0 + 0.5 * x1) / 8.0 +
                I[I_B2P] * 3.0 * (0.5 + x1) / 8.0);
    d[5] = s * (I[I_B1] * t1 / 8.0 + I[I_B1P] * t1 / 4.0 + I[I_B2] * 3.0 * t1 / 16.0 + I[I_B2P] * 3.0 * t1 / 8.0);
    d[2] = s * (I[I_B1] * (1.0 + 0.5 * x2) / 4.0 - I[I_B1P] * (0.5 + x2) / 4.0 - I[I_B2] * (1.0 + 0.5 * x2) / 8.0 +
                I[I_B2P] * (0.5 + x2) / 8.0);
    d[6] = s * (I[I_B1] * t2 / 8.0 - I[I_B1P] * t2 / 4.0 - I[I_B2] * t2 / 16.0 + I[I_B2P] * t2 / 8.0);
    // b3 = t3 (1 + x3/2) / 4, b3' = t3 (1/2 + x3) / 4
    d[3] = s * (I[I_B3] * (1.0 + 0.5 * x3) / 4.0 + I[I_B3P] * (0.5 + x3) / 4.0);
    d[7] = s * (I[I_B3] * t3 / 8.0 + I[I_B3P] * t3 / 4.0);
    d[8] = s * I[I_ALPHA];
    d[9] = s * 0.5 * I[I_B4];  // b4 = t4 / 2
  }
}

}  // namespace

extern "C" int hfb_energy_bwd_nblocks(hfb_ctx* c) {
  if (!c) return HFB_E_NOCTX;
  const int64_t n = (int64_t)c->nx * c->ny * c->nz;
  int nb = (int)((n + 255) / 256);
  return nb > 512 ? 512 : nb;
}

// forces: nbatch host structs (the b-couplings of each set, as for
// hfb_integ_energy); tx: device (nbatch, 10) raw couplings (t0..t3, x0..x3,
// power, t4); gout: device (nbatch) upstream gradients.  Density gradients
// (summed over the batch with gout) go to the d_* fields, each may be null;
// coupling gradients to d_coupling (device, nbatch x 10) when non-null.
// part: hfb_energy_bwd_nblocks() * 72 doubles.
extern "C" int hfb_energy_bwd(hfb_ctx* c, const hfb_force* forces, const double* tx, const double* gout, int nbatch,
                              const double* rho, const double* tau, const double* sodens, const double* current,
                              const double* sdens, const double* divj, const double* lap, const double* grad,
                              const double* curls, const double* curlj, const double* wcoul, double* d_rho,
                              double* d_tau, double* d_sodens, double* d_current, double* d_sdens, double* d_coupling,
                              double* part, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!forces || !gout || !rho || !tau || !sodens || !current || !sdens || !divj || !lap || !grad || !curls ||
      !curlj || !part)
    return HFB_E_ARG;
  if (nbatch < 1 || nbatch > MAXB) return HFB_E_SIZE;
  if (d_coupling && !tx) return HFB_E_ARG;
  AdjArgs a{};
  a.nb = nbatch;
  a.gout = gout;
  for (int b = 0; b < nbatch; ++b) {
    a.f[b] = forces[b];
    if (forces[b].tcoul && !wcoul) return HFB_E_ARG;
  }
  a.nvox = (int64_t)c->nx * c->ny * c->nz;
  a.wxyz = c->dx * c->dy * c->dz;
  a.rho = rho;
  a.tau = tau;
  a.sodens = sodens;
  a.cur = current;
  a.sdens = sdens;
  a.divj = divj;
  a.lap = lap;
  a.grad = grad;
  a.curls = curls;
  a.curlj = curlj;
  a.wcoul = wcoul;
  a.d_rho = d_rho;
  a.d_tau = d_tau;
  a.d_sodens = d_sodens;
  a.d_cur = d_current;
  a.d_sdens = d_sdens;
  a.part = part;
  const int nblk = hfb_energy_bwd_nblocks(c);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(energy_bwd_kernel, dim3(nblk), dim3(256), 0, st, a);
  int rc = launch_status();
  if (rc != HFB_OK || !d_coupling) return rc;
  hipLaunchKernelGGL(energy_bwd_finish_kernel, dim3(nbatch), dim3(256), 0, st, part, nblk, a.wxyz, tx, gout,
                     d_coupling);
  return launch_status();
}
