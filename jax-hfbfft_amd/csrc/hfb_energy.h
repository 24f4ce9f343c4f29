// The energy functional's per-block partial layout (energies.py:90-306) and
// the one-workgroup bodies that finish an SCF iteration's scalars: ecorrp
// and the v_pair maxima, the energy integrals, sum_energy.  Shared by the
// energy tail launch and by the h.psi launches, whose extra workgroups can
// run the previous iteration's tail.  The finishing bodies take workgroups
// of at most 256 threads, energy_tail_body up to 1024.
#pragma once
#include <math.h>

#include "hfb_ctx.h"

namespace hfb {

constexpr int NES = 20;
enum {
  S_EHF0, S_CRHO0, S_CRHO1, S_EHF3, S_EHF2, S_CDRHO0, S_CDRHO1, S_CTAU0, S_CTAU1, S_EHF1,
  S_EHFLS, S_CDJ0, S_CDJ1, S_LSODD, S_EHFC, S_ECORC, S_EHFT, S_ECORRP_UNUSED, S_PAD0, S_PAD1
};

struct EnArgs {
  hfb_force f;
  int64_t nvox;
  const double *rho, *tau, *sdens;
  const double *divj, *lap, *curlj;
  const double* wcoul;
  double* part;  // [block][NES]
};

// energies.py:157-167 index current[:,:,:,0,0] .. current[:,:,:,2,1]: the
// (y, z) = (0..2, 0..1) corner of every (iq, component, x) -- kept as written.
// Item it < 18 nx of the two C_j sums reads p[0], p[1] at this address.
__device__ __forceinline__ const double* cj_item(const double* cur, int it, int nx, int ny, int nz) {
  const int y = it % 3;
  const int x = (it / 3) % nx;
  const int ic = it / (3 * nx);  // iq*3 + component
  return cur + (int64_t)ic * nx * ny * nz + (int64_t)x * ny * nz + (int64_t)y * nz;
}
__device__ __forceinline__ void cj_add(double a0, double a1, double* cj) {
  cj[0] += (a0 + a1) * (a0 + a1);
  cj[1] += (-a0 + a1) * (-a0 + a1);
}

// Adds this thread's share of the two C_j sums to cj.
__device__ __forceinline__ void cj_terms(const double* __restrict__ cur, int nx, int ny, int nz, double* cj) {
  for (int it = threadIdx.x; it < 2 * 3 * nx * 3; it += blockDim.x) {
    const double* p = cj_item(cur, it, nx, ny, nz);
    cj_add(p[0], p[1], cj);
  }
}

// What sum_energy reads of the integrals' outputs.
struct EnKeep {
  double e3corr, ecorc, ezpe;
};

// The integrals (energies.py:90-306) from the summed partials s[NES], the
// C_j sums and the pairing energies; one thread.
__device__ inline EnKeep energy_outputs(const hfb_force& f, const double* s, double cj0, double cj1, double wxyz,
                                        double epair, double* __restrict__ E) {
  const double ehf0 = wxyz * s[S_EHF0], ehf1 = wxyz * s[S_EHF1], ehf2 = wxyz * s[S_EHF2], ehf3 = wxyz * s[S_EHF3];
  const double ehflsodd = wxyz * s[S_LSODD];
  const double ehfls = wxyz * s[S_EHFLS] + ehflsodd;
  const double ehfc = f.tcoul ? wxyz * s[S_EHFC] : 0.0;
  const double ecorc = f.tcoul ? wxyz * s[S_ECORC] : 0.0;
  const double ehft = wxyz * s[S_EHFT];
  const double ezpe = f.zpe == 1 ? 17.3 / pow(f.mass_number, 0.2) : 0.0;
  const double e3corr = -f.power * ehf3 / 2.0;
  E[HFB_EN_EHFT] = ehft;
  E[HFB_EN_EHF0] = ehf0;
  E[HFB_EN_EHF1] = ehf1;
  E[HFB_EN_EHF2] = ehf2;
  E[HFB_EN_EHF3] = ehf3;
  E[HFB_EN_EHFLS] = ehfls;
  E[HFB_EN_EHFLSODD] = ehflsodd;
  E[HFB_EN_EHFC] = ehfc;
  E[HFB_EN_ECORC] = ecorc;
  E[HFB_EN_E3CORR] = e3corr;
  E[HFB_EN_EZPE] = ezpe;
  E[HFB_EN_CRHO0] = wxyz * s[S_CRHO0];
  E[HFB_EN_CRHO1] = wxyz * s[S_CRHO1];
  E[HFB_EN_CDRHO0] = wxyz * s[S_CDRHO0];
  E[HFB_EN_CDRHO1] = wxyz * s[S_CDRHO1];
  E[HFB_EN_CTAU0] = wxyz * s[S_CTAU0];
  E[HFB_EN_CTAU1] = wxyz * s[S_CTAU1];
  E[HFB_EN_CDJ0] = wxyz * s[S_CDJ0];
  E[HFB_EN_CDJ1] = wxyz * s[S_CDJ1];
  E[HFB_EN_CJ0] = -wxyz * f.Ctau0 * cj0;
  E[HFB_EN_CJ1] = -wxyz * f.Ctau1 * cj1;
  E[HFB_EN_EHFINT] = ehft + ehf0 + ehf1 + ehf2 + ehf3 + ehfls + ehfc - epair - ezpe;
  return EnKeep{e3corr, ecorc, ezpe};
}

// sum_energy (energies.py:309-371) from the state sums v[8] (w(kin + spe),
// w kin, w orb[3], w spin[3]) and the integrals' e3corr, ecorc, ezpe; one
// thread.
__device__ inline void sum_energy_outputs(const double* v, const EnKeep& k, double ecorrp, double epair, double* E) {
  E[HFB_EN_EHF] = v[0] / 2.0 + k.e3corr + ecorrp + k.ecorc - epair - k.ezpe;
  E[HFB_EN_TKE] = v[1];
  for (int c = 0; c < 3; ++c) {
    E[HFB_EN_ORB + c] = v[2 + c];
    E[HFB_EN_SPIN + c] = v[5 + c];
    E[HFB_EN_JTOT + c] = v[2 + c] + v[5 + c];
  }
}

// This thread's share of the state sums of sum_energy into v[8].
__device__ __forceinline__ void sum_energy_terms(int ns, const double* wocc, const double* wstates,
                                                 const double* kin, const double* spe, const double* orb,
                                                 const double* spin, double* v) {
  for (int s = threadIdx.x; s < ns; s += blockDim.x) {
    const double w = wocc[s] * wstates[s];
    v[0] += w * (kin[s] + spe[s]);
    v[1] += w * kin[s];
    for (int c = 0; c < 3; ++c) {
      v[2 + c] += w * orb[3 * s + c];
      v[5 + c] += w * spin[3 * s + c];
    }
  }
}

// Maximum of (m0, m1) over the workgroup into out[0..1] for every thread.
__device__ __forceinline__ void block_max2(double m0, double m1, double (*vred)[2], double* out) {
  for (int o = 32; o > 0; o >>= 1) {
    m0 = fmax(m0, __shfl_xor(m0, o, 64));
    m1 = fmax(m1, __shfl_xor(m1, o, 64));
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    vred[wid][0] = m0;
    vred[wid][1] = m1;
  }
  __syncthreads();
  out[0] = vred[0][0];
  out[1] = vred[0][1];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
    out[0] = fmax(out[0], vred[w][0]);
    out[1] = fmax(out[1], vred[w][1]);
  }
}

// The integrals from energy_kernel's block partials (fixed order) into E
// (energies.py:90-306); one workgroup.
__device__ inline void energy_finish_body(const EnArgs& a, int nblk, double wxyz, const double* __restrict__ cur,
                                          int nx, int ny, int nz, double* __restrict__ E) {
  __shared__ double red[4 * (NES + 2)];
  double s[NES + 2];
  for (int j = 0; j < NES + 2; ++j) s[j] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
    for (int j = 0; j < NES; ++j) s[j] += a.part[(int64_t)b * NES + j];
  cj_terms(cur, nx, ny, nz, s + NES);
  block_sum<NES + 2>(s, red);
  if (threadIdx.x == 0) energy_outputs(a.f, s, s[NES], s[NES + 1], wxyz, E[HFB_EN_EPAIR] + E[HFB_EN_EPAIR + 1], E);
}

// ecorrp from the block partials (meanfield.py:360-363) and the maxima of
// v_pair per isospin into E[HFB_EN_VMAX..+1].
__device__ inline void ecorrp_finish_body(const double* part, int nblk, double wxyz, int ipair, double* E) {
  __shared__ double red[4];
  __shared__ double vred[16][2];
  double s[1] = {0.0};
  double m0 = -INFINITY, m1 = -INFINITY;
  for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
    s[0] += part[b];
    m0 = fmax(m0, part[nblk + b]);
    m1 = fmax(m1, part[2 * nblk + b]);
  }
  block_sum<1>(s, red);
  double vm[2];
  block_max2(m0, m1, vred, vm);
  if (threadIdx.x == 0) {
    E[HFB_EN_ECORRP] = ipair == 6 ? -s[0] * wxyz / 2.0 : 0.0;
    E[HFB_EN_VMAX] = vm[0];
    E[HFB_EN_VMAX + 1] = vm[1];
  }
}

// sum_energy (energies.py:309-371)
__device__ inline void sum_energy_body(int ns, const double* wocc, const double* wstates, const double* kin,
                                       const double* spe, const double* orb, const double* spin, double* E) {
  __shared__ double red[16 * 8];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  sum_energy_terms(ns, wocc, wstates, kin, spe, orb, spin, v);
  block_sum<8>(v, red);
  if (threadIdx.x == 0)
    sum_energy_outputs(v, EnKeep{E[HFB_EN_E3CORR], E[HFB_EN_ECORC], E[HFB_EN_EZPE]}, E[HFB_EN_ECORRP],
                       E[HFB_EN_EPAIR] + E[HFB_EN_EPAIR + 1], E);
}

// Everything the tail reads, with the grid constants of the context.
struct TailArgs {
  EnArgs en;
  const double* sky_part;
  int nblk;
  double wxyz;
  const double* cur;
  int nx, ny, nz;
  int ns;
  const double *wocc, *wstates, *kin, *spe, *orb, *spin;
  double* E;
};

// The v_pair maxima per isospin from the skyrme partials (the values the
// tail writes to E[HFB_EN_VMAX..+1]) into vm[0..1] of every thread; at
// most 64 * MW threads.
template <int MW = 16>
__device__ inline void vpair_max_of_parts(const TailArgs& t, double* vm) {
  __shared__ double vred[MW][2];
  double m0 = -INFINITY, m1 = -INFINITY;
  for (int b = threadIdx.x; b < t.nblk; b += blockDim.x) {
    m0 = fmax(m0, t.sky_part[t.nblk + b]);
    m1 = fmax(m1, t.sky_part[2 * t.nblk + b]);
  }
  block_max2(m0, m1, vred, vm);
}

// ecorrp and the v_pair maxima, the integrals, then sum_energy, in one
// pass: every load is issued before the first reduction, the sums share one
// block_sum and nothing written is read back, so the workgroup waits on
// memory once.  Returns the v_pair maxima in vm[0..1] of every thread.  At
// most 64 * MW threads.
template <int MW = 16>
__device__ inline void energy_tail_body(const TailArgs& t, double* vm) {
  constexpr int NV = 1 + NES + 2 + 8;  // ecorrp, integrals, C_j, state sums
  __shared__ double red[MW * NV];
  __shared__ double vred[MW][2];
  double* E = t.E;
  const int tid = threadIdx.x, bd = blockDim.x;
  // every load of this thread's first partial block, first CJU C_j items
  // and first state is issued before any is used; what lies beyond (more
  // blocks, items or states than that) is added by the loops after
  constexpr int CJU = 4;
  const int ncj = 2 * 3 * t.nx * 3;
  const double ep0 = E[HFB_EN_EPAIR], ep1 = E[HFB_EN_EPAIR + 1];
  double sk0 = 0.0, m0 = -INFINITY, m1 = -INFINITY, en[NES];
#pragma unroll
  for (int j = 0; j < NES; ++j) en[j] = 0.0;
  if (tid < t.nblk) {
    sk0 = t.sky_part[tid];
    m0 = t.sky_part[t.nblk + tid];
    m1 = t.sky_part[2 * t.nblk + tid];
#pragma unroll
    for (int j = 0; j < NES; ++j) en[j] = t.en.part[(int64_t)tid * NES + j];
  }
  double ca[CJU], cb[CJU];
#pragma unroll
  for (int u = 0; u < CJU; ++u) {
    const int it = tid + u * bd;
    ca[u] = cb[u] = 0.0;
    if (it < ncj) {
      const double* p = cj_item(t.cur, it, t.nx, t.ny, t.nz);
      ca[u] = p[0];
      cb[u] = p[1];
    }
  }
  double so[10];
#pragma unroll
  for (int j = 0; j < 10; ++j) so[j] = 0.0;
  if (tid < t.ns) {
    so[0] = t.wocc[tid];
    so[1] = t.wstates[tid];
    so[2] = t.kin[tid];
    so[3] = t.spe[tid];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      so[4 + c] = t.orb[3 * tid + c];
      so[7 + c] = t.spin[3 * tid + c];
    }
  }
  double v[NV];
  v[0] = sk0;
#pragma unroll
  for (int j = 0; j < NES; ++j) v[1 + j] = en[j];
  v[1 + NES] = v[2 + NES] = 0.0;
#pragma unroll
  for (int u = 0; u < CJU; ++u) cj_add(ca[u], cb[u], v + 1 + NES);
  {
    const double w = so[0] * so[1];
    v[3 + NES] = w * (so[2] + so[3]);
    v[4 + NES] = w * so[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      v[5 + NES + c] = w * so[4 + c];
      v[8 + NES + c] = w * so[7 + c];
    }
  }
  for (int b = tid + bd; b < t.nblk; b += bd) {
    v[0] += t.sky_part[b];
    m0 = fmax(m0, t.sky_part[t.nblk + b]);
    m1 = fmax(m1, t.sky_part[2 * t.nblk + b]);
#pragma unroll
    for (int j = 0; j < NES; ++j) v[1 + j] += t.en.part[(int64_t)b * NES + j];
  }
  for (int it = tid + CJU * bd; it < ncj; it += bd) {
    const double* p = cj_item(t.cur, it, t.nx, t.ny, t.nz);
    cj_add(p[0], p[1], v + 1 + NES);
  }
  if (t.ns > bd) {
    double rest[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = tid + bd; s < t.ns; s += bd) {
      const double w = t.wocc[s] * t.wstates[s];
      rest[0] += w * (t.kin[s] + t.spe[s]);
      rest[1] += w * t.kin[s];
      for (int c = 0; c < 3; ++c) {
        rest[2 + c] += w * t.orb[3 * s + c];
        rest[5 + c] += w * t.spin[3 * s + c];
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) v[3 + NES + j] += rest[j];
  }
  const double epair = ep0 + ep1;
  block_sum<NV>(v, red);
  block_max2(m0, m1, vred, vm);
  if (threadIdx.x == 0) {
    const double ecorrp = t.en.f.ipair == 6 ? -v[0] * t.wxyz / 2.0 : 0.0;
    E[HFB_EN_ECORRP] = ecorrp;
    E[HFB_EN_VMAX] = vm[0];
    E[HFB_EN_VMAX + 1] = vm[1];
    const EnKeep k = energy_outputs(t.en.f, v + 1, v[1 + NES], v[2 + NES], t.wxyz, epair, E);
    sum_energy_outputs(v + 3 + NES, k, ecorrp, epair, E);
  }
}

// TailArgs from the public record; false if an argument is missing.
inline bool tail_args(const hfb_ctx* c, const hfb_tail_job& j, TailArgs& t) {
  if (!j.force || !j.sky_part || !j.en_part || !j.current || j.nstates <= 0 || !j.wocc || !j.wstates ||
      !j.sp_kinetic || !j.sp_energy || !j.sp_orbital || !j.sp_spin || !j.E)
    return false;
  t = TailArgs{};
  t.en.f = *j.force;
  t.en.nvox = (int64_t)c->nx * c->ny * c->nz;
  t.en.part = (double*)j.en_part;
  t.sky_part = j.sky_part;
  t.nblk = hfb_skyrme_nblocks(const_cast<hfb_ctx*>(c));
  t.wxyz = c->dx * c->dy * c->dz;
  t.cur = j.current;
  t.nx = c->nx;
  t.ny = c->ny;
  t.nz = c->nz;
  t.ns = j.nstates;
  t.wocc = j.wocc;
  t.wstates = j.wstates;
  t.kin = j.sp_kinetic;
  t.spe = j.sp_energy;
  t.orb = j.sp_orbital;
  t.spin = j.sp_spin;
  t.E = j.E;
  return true;
}

}  // namespace hfb
