// Scalar jobs of the gradient step (static.py:93-259) as workgroup bodies:
// run by their own one-workgroup launch, or by one extra workgroup of the
// h.psi x pass (weights) and of the Gram launch (finish), which do not read
// what these jobs write.  Any workgroup size that is a multiple of 64 (at
// most 1024) works.
#pragma once
#include "hfb_ctx.h"
#include "hfb_energy.h"

namespace hfb {

inline bool weights_job_ok(const hfb_weights_job& j) {
  return j.nstates > 0 && j.wocc && j.wstates && j.wguv && j.pairwg && (j.v_pair || j.vpair_max) && j.w && j.wuv &&
         j.pw && j.pshift;
}

inline bool finish_job_ok(const hfb_finish_job& j) {
  return j.nstates > 0 && j.hsum && j.sp_energy && j.wocc && j.wstates && j.sp_efluct1 && j.sp_efluct2 && j.acc &&
         j.hsum_ntile >= 0 && j.fsum_ntile >= 0;
}

// The weights of state s from its occupation inputs and the v_pair maxima
// per isospin vmax[2].
__device__ __forceinline__ void weights_state(const hfb_weights_job& job, int s, double wocc, double wstates,
                                              double wguv, double pairwg, const double* vmax) {
  const int iq = s >= job.npsi0 ? 1 : 0;
  if (job.tbcs) {
    job.w[s] = 1.0;
    job.wuv[s] = 0.0;
    job.pw[s] = 1.0;
    job.pshift[s] = 0.0;  // laplace(..., 1.0, 0.0, ps1, 0.0, e0dmp)
  } else {
    job.w[s] = wstates * wocc;
    job.wuv[s] = wstates * wguv * pairwg;
    job.pw[s] = fmax(wocc, 0.1);
    job.pshift[s] = 0.5 * fmax(wguv * pairwg, 0.1) * vmax[iq];
  }
}

// Per-state weights of the gradient step (static.py:152-170, 125-136):
// hpsi weights (w, wuv) and the preconditioner's (max(w,.1), shift).
__device__ __forceinline__ void grstep_weights_body(const hfb_weights_job& job) {
  const int ns = job.nstates, tbcs = job.tbcs;
  const double* __restrict__ wocc = job.wocc;
  const double* __restrict__ wstates = job.wstates;
  const double* __restrict__ wguv = job.wguv;
  const double* __restrict__ pairwg = job.pairwg;
  const double* __restrict__ vpair = job.v_pair;
  const int64_t nvox = job.nvox;
  const double* __restrict__ vpair_max = job.vpair_max;
  __shared__ double vmax_s[2];
  __shared__ double red[16];
  for (int iq = 0; iq < 2; ++iq) {
    double m = -INFINITY;
    if (!tbcs && vpair_max) {
      if (threadIdx.x == 0) m = vpair_max[iq];
    } else if (!tbcs) {
      // eight loads in flight per thread (four independent max chains); the
      // field offset iq * nvox keeps 16-byte alignment when nvox is even
      const double* v = vpair + iq * nvox;
      const int64_t bd = blockDim.x;
      int64_t i = threadIdx.x;
      if ((nvox & 1) == 0) {
        const double2* v2 = reinterpret_cast<const double2*>(v);
        const int64_t n2 = nvox / 2;
        double m0 = -INFINITY, m1 = -INFINITY, m2 = -INFINITY, m3 = -INFINITY;
        for (; i + 3 * bd < n2; i += 4 * bd) {
          const double2 a = v2[i], b = v2[i + bd], c = v2[i + 2 * bd], d = v2[i + 3 * bd];
          m0 = fmax(m0, fmax(a.x, a.y));
          m1 = fmax(m1, fmax(b.x, b.y));
          m2 = fmax(m2, fmax(c.x, c.y));
          m3 = fmax(m3, fmax(d.x, d.y));
        }
        for (; i < n2; i += bd) m0 = fmax(m0, fmax(v2[i].x, v2[i].y));
        m = fmax(fmax(m0, m1), fmax(m2, m3));
      } else {
        for (; i < nvox; i += bd) m = fmax(m, v[i]);
      }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    if (lane == 0) red[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
      double mm = -INFINITY;
      for (int j = 0; j < (int)(blockDim.x + 63) / 64; ++j) mm = fmax(mm, red[j]);
      vmax_s[iq] = mm;
    }
    __syncthreads();
  }
  for (int s = threadIdx.x; s < ns; s += blockDim.x)
    weights_state(job, s, wocc[s], wstates[s], wguv[s], pairwg[s], vmax_s);
}

// Bookkeeping after the step (static.py:173, 211, 247-257).
// hsum / fsum: per-state totals (nt == 0) or tile partials (nt > 0; fsum's
// partials then still carry no wxyz factor).
__device__ __forceinline__ void grstep_finish_body(const hfb_finish_job& job) {
  const int ns = job.nstates, hsum_nt = job.hsum_ntile, fsum_nt = job.fsum_ntile;
  const double wxyz = job.wxyz;
  const double* __restrict__ hsum = job.hsum;
  double* __restrict__ sp_energy = job.sp_energy;
  const double* __restrict__ wocc = job.wocc;
  const double* __restrict__ wstates = job.wstates;
  const double* __restrict__ fsum = job.fsum;
  double* __restrict__ efl1 = job.sp_efluct1;
  double* __restrict__ efl2 = job.sp_efluct2;
  double* __restrict__ acc = job.acc;
  __shared__ double red[16 * 2];
  const double fsc = fsum_nt > 0 ? wxyz : 1.0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  double dsum = 0.0;  // lane 0: sum of wocc*wstates*denerg over this wavefront's states
  // one wavefront per state: the lanes take the tile partials, a butterfly
  // adds them (the same total in every lane)
  for (int s = wave; s < ns; s += nwave) {
    double t[3] = {0.0, 0.0, 0.0};
    if (hsum_nt > 0) {
      for (int k = lane; k < hsum_nt; k += 64) t[0] += hsum[((int64_t)s * hsum_nt + k) * 2];
    } else if (lane == 0) {
      t[0] = hsum[2 * s];
    }
    if (fsum) {
      if (fsum_nt > 0) {
        for (int k = lane; k < fsum_nt; k += 64) {
          t[1] += fsum[((int64_t)s * fsum_nt + k) * 2];
          t[2] += fsum[((int64_t)s * fsum_nt + k) * 2 + 1];
        }
      } else if (lane == 0) {
        t[1] = fsum[2 * s];
        t[2] = fsum[2 * s + 1];
      }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      t[j] += lane_xor<1>(t[j]);
      t[j] += lane_xor<2>(t[j]);
      t[j] += lane_xor<4>(t[j]);
      t[j] += lane_xor<8>(t[j]);
      t[j] += lane_xor<16>(t[j]);
      t[j] += __shfl_xor(t[j], 32, 64);
    }
    if (lane == 0) {
      const double spe_old = sp_energy[s];
      const double spe_new = wxyz * t[0];
      dsum += wocc[s] * wstates[s] * ((spe_old - spe_new) / fabs(spe_new));
      if (fsum) {
        efl1[s] = sqrt(fsc * t[1]);
        efl2[s] = sqrt(fsc * t[2]);
      }
      sp_energy[s] = spe_new;
    }
  }
  __syncthreads();
  // sums over the states (static.py:211, 255-257)
  double v[2] = {0.0, dsum};
  for (int s = threadIdx.x; s < ns; s += blockDim.x) v[0] += efl1[s];
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    acc[0] += v[0];
    acc[1] += v[1];
  }
}

// The jobs an h.psi launch runs in its extra first row of workgroups: the
// step weights (w.nstates > 0) and/or the energy tail of the previous
// iteration (tail; hfb_weights_job.tail).  vm_part: the weights take the
// v_pair maxima from the tail's skyrme partials (when the tail runs in
// another launch of the same call).
struct WJob {
  hfb_weights_job w;
  TailArgs t;
  int tail;
  int vm_part;
};

// A WJob from the public record (null: no job); false if an argument is
// missing.
inline bool wjob_of(const hfb_ctx* c, const hfb_weights_job* job, WJob& out) {
  out = WJob{};
  if (!job) return true;
  out.w = *job;
  out.w.tail = nullptr;
  if (!weights_job_ok(out.w)) return false;
  if (job->tail) {
    if (!tail_args(c, *job->tail, out.t)) return false;
    out.tail = 1;
  }
  return true;
}

// Rows of workgroups a launch adds in front of its states for the job.
__host__ __device__ inline int job_rows(const WJob& j) { return (j.w.nstates > 0 || j.tail) ? 1 : 0; }

// The weights part of a job (no tail) and the tail part (no weights), for
// calls whose first launch runs the weights and a later one the tail.
inline WJob weights_part(const WJob& j) {
  WJob o = j;
  o.vm_part = j.tail;
  o.tail = 0;
  return o;
}
inline WJob tail_part(const WJob& j) {
  WJob o{};
  o.t = j.t;
  o.tail = j.tail;
  return o;
}

// The state inputs of the weights of state s0 (zeros past the last state).
struct WIn {
  double wo, wst, wg, pg;
};
__device__ __forceinline__ WIn weights_in(const hfb_weights_job& w, int s) {
  if (s >= w.nstates) return WIn{0.0, 0.0, 0.0, 0.0};
  return WIn{w.wocc[s], w.wstates[s], w.wguv[s], w.pairwg[s]};
}

// The job of one launch, run by one workgroup of at most 64 * MW threads.
// With the tail or vm_part the weights take the v_pair maxima from the
// skyrme partials (the values the tail writes to E[HFB_EN_VMAX]) and their
// inputs are loaded first.
template <int MW>
__device__ __forceinline__ void job_body(const WJob& j) {
  const hfb_weights_job& w = j.w;
  if (w.nstates > 0 && !j.tail && !j.vm_part) {
    grstep_weights_body(w);
    return;
  }
  double vm[2];
  if (w.nstates <= 0) {  // the tail alone
    energy_tail_body<MW>(j.t, vm);
    return;
  }
  const int s0 = threadIdx.x;
  WIn in = weights_in(w, s0);
  if (j.tail)
    energy_tail_body<MW>(j.t, vm);
  else
    vpair_max_of_parts<MW>(j.t, vm);
  for (int s = s0; s < w.nstates; s += blockDim.x) {
    if (s != s0) in = weights_in(w, s);
    weights_state(w, s, in.wo, in.wst, in.wg, in.pg, vm);
  }
}

}  // namespace hfb
