// Mean-field potentials from the densities (meanfield.py:163-391) and the
// energy-functional integrals (energies.py:90-371), FP64.
//
// The reference's derivatives of real fields are dense n x n Fourier
// matrices applied along one axis (der1 = sder, der2 = sder2; grids.py:36-61,
// einsum forms 'ij,jkl->ikl', 'jl,ilk->ijk', 'kl,ijl->ijk').  They are
// applied here as axis stencils: each output point sums n products along the
// axis, with neighbouring threads on consecutive z so the reads coalesce.
// This work is O(n^4) per field, independent of the number of states.
#include <math.h>

#include "hfb_ctx.h"
#include "hfb_energy.h"
#include "hfb_poisson.h"

using namespace hfb;

namespace {

// One derived field = sum of up to three axis derivatives of source fields.
struct DerivJob {
  const double* src[3];
  int axis[3];
  int order[3];  // 1: der1, 2: der2
  double coef[3];
  int nterm;
  double* dst;
};

constexpr int MAXJOB = 24;
constexpr int MAXTERM = 3 * MAXJOB;
struct DerivJobs {
  DerivJob j[MAXJOB];
  int njob;
  // every term of every job, in job order: job index, term index, and the
  // slot of the term in the workspace (-1: single-term job, written to dst)
  int tjob[MAXTERM], tidx[MAXTERM], tslot[MAXTERM];
  int nterm_all;
  int slot0[MAXJOB];  // workspace slot of term 0 of a multi-term job
};

__global__ void __launch_bounds__(256) deriv_kernel(DerivJobs jobs, int nx, int ny, int nz, const double* __restrict__ d1x,
                                                    const double* __restrict__ d1y, const double* __restrict__ d1z,
                                                    const double* __restrict__ d2x, const double* __restrict__ d2y,
                                                    const double* __restrict__ d2z) {
  const DerivJob& jb = jobs.j[blockIdx.y];
  const int64_t nvox = (int64_t)nx * ny * nz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nvox; e += (int64_t)gridDim.x * blockDim.x) {
    const int x = e / ((int64_t)ny * nz);
    const int y = (e / nz) % ny;
    const int z = e % nz;
    double acc = 0.0;
    for (int t = 0; t < jb.nterm; ++t) {
      const double* s = jb.src[t];
      const int ax = jb.axis[t];
      const double* D;
      int n, i;
      int64_t base, stride;
      if (ax == 0) {
        D = jb.order[t] == 1 ? d1x : d2x;
        n = nx;
        i = x;
        base = (int64_t)y * nz + z;
        stride = (int64_t)ny * nz;
      } else if (ax == 1) {
        D = jb.order[t] == 1 ? d1y : d2y;
        n = ny;
        i = y;
        base = (int64_t)x * ny * nz + z;
        stride = nz;
      } else {
        D = jb.order[t] == 1 ? d1z : d2z;
        n = nz;
        i = z;
        base = ((int64_t)x * ny + y) * nz;
        stride = 1;
      }
      double v = 0.0;
      for (int l = 0; l < n; ++l) v += D[i * n + l] * s[base + l * stride];
      acc += jb.coef[t] * v;
    }
    jb.dst[e] = acc;
  }
}

// Tiled form for cubic grids: a workgroup stages 64 pencils of one source
// field along the term's axis in LDS and each wavefront produces n/4 output
// rows for all 64 pencils (out = D s, D row-uniform across the wavefront, so
// its entries come through the scalar cache).  One launch per term index;
// later terms accumulate into dst in stream order.
constexpr int DT_P = 64;          // pencils per tile
constexpr int DT_LD = DT_P + 1;   // LDS row pitch (doubles)

struct DerivMats {
  const double* d1[3];
  const double* d2[3];
};

template <int N>
__device__ __forceinline__ int64_t pencil_off(int ax, int k, int gp) {
  if (ax == 0) return (int64_t)k * N * N + gp;
  if (ax == 1) return (int64_t)(gp / N) * N * N + (int64_t)k * N + gp % N;
  return (int64_t)gp * N + k;
}

// One launch for every term of every job: block (tile, term) computes
// coef * D src along the term's axis for 64 pencils and writes it to its
// workspace slot (or to dst for a single-term job); deriv_sum_kernel then
// adds the terms of each job in term order.
// Tile tx of term tg by 256 threads (index tid); sS: N * DT_LD doubles of
// LDS.  An inactive call only takes part in the workgroup barriers.
template <int N>
__device__ __forceinline__ void deriv_tile_body(const DerivJobs& jobs, const DerivMats& dm, double* __restrict__ work,
                                                int tg, int tx, bool active, double* sS, int tid) {
  const DerivJob& jb = jobs.j[active ? jobs.tjob[tg] : 0];
  const int term = active ? jobs.tidx[tg] : 0;
  const int slot = active ? jobs.tslot[tg] : -1;
  const int ax = jb.axis[term];
  const double* __restrict__ src = jb.src[term];
  const double* __restrict__ D = jb.order[term] == 1 ? dm.d1[ax] : dm.d2[ax];
  const double coef = jb.coef[term];
  double* __restrict__ dst = slot < 0 ? jb.dst : work + (int64_t)slot * N * N * N;
  const int p0 = tx * DT_P;
  // element (k, p) of the tile; consecutive threads walk the contiguous direction
  auto kp = [&](int e, int& k, int& p) {
    if (ax == 2) {
      k = e % N;
      p = e / N;
    } else {
      p = e % DT_P;
      k = e / DT_P;
    }
  };
  if (active)
    for (int e = tid; e < N * DT_P; e += 256) {
      int k, p;
      kp(e, k, p);
      sS[k * DT_LD + p] = src[pencil_off<N>(ax, k, p0 + p)];
    }
  __syncthreads();
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int p = tid & 63;
  constexpr int R = N / 4;
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.0;
  if (active) {
#pragma unroll
    for (int l = 0; l < N; ++l) {
      const double sv = sS[l * DT_LD + p];
#pragma unroll
      for (int r = 0; r < R; ++r) acc[r] = fma(D[(w + 4 * r) * N + l], sv, acc[r]);
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int r = 0; r < R; ++r) sS[(w + 4 * r) * DT_LD + p] = acc[r];
  }
  __syncthreads();
  if (active)
    for (int e = tid; e < N * DT_P; e += 256) {
      int k, pp;
      kp(e, k, pp);
      dst[pencil_off<N>(ax, k, p0 + pp)] = coef * sS[k * DT_LD + pp];
    }
}

template <int N>
__global__ void __launch_bounds__(256) deriv_tile_kernel(DerivJobs jobs, DerivMats dm, double* __restrict__ work) {
  __shared__ double sS[N * DT_LD];
  deriv_tile_body<N>(jobs, dm, work, blockIdx.y, blockIdx.x, true, sS, threadIdx.x);
}

// dst = term 0 + term 1 (+ term 2) for every multi-term job, in that order:
// job jo, workgroup bx of gx (256 threads).
__device__ __forceinline__ void deriv_sum_body(const DerivJobs& jobs, const double* __restrict__ work, int64_t nvox,
                                               int bx, int gx, int jo) {
  const DerivJob& jb = jobs.j[jo];
  if (jb.nterm < 2) return;
  const double* w0 = work + (int64_t)jobs.slot0[jo] * nvox;
  for (int64_t e = bx * (int64_t)256 + threadIdx.x; e < nvox; e += (int64_t)gx * 256) {
    double v = w0[e] + w0[nvox + e];
    if (jb.nterm == 3) v += w0[2 * nvox + e];
    jb.dst[e] = v;
  }
}

__global__ void __launch_bounds__(256) deriv_sum_kernel(DerivJobs jobs, const double* __restrict__ work,
                                                        int64_t nvox) {
  deriv_sum_body(jobs, work, nvox, blockIdx.x, gridDim.x, blockIdx.y);
}

struct SkyArgs {
  hfb_force f;
  int64_t nvox;
  const double *rho, *tau, *chi, *cur, *sdens;  // densities
  const double *divj, *lap, *grad, *curls, *curlj;  // derived, [2] / [2][3]
  const double* wcoul;
  double *upot, *bmass, *wlspot, *aq, *spot, *vpair;  // FP64 out
  float *upot32, *bmass32, *wlspot32, *aq32, *spot32, *vpair32;  // optional FP32 copies
  double* epart;  // per-block partials: [0, nb) sum rho_tot * rearr (DDDI), [nb, 3 nb) max v_pair per isospin
};

// (workgroup vb of vg)
__device__ __forceinline__ void skyrme_body(const SkyArgs& a, int vb, int vg) {
  const hfb_force& f = a.f;
  const int64_t n = a.nvox;
  __shared__ double red[4];
  __shared__ double vred[4][2];
  double ep[1] = {0.0};
  double vm[2] = {-INFINITY, -INFINITY};
  for (int64_t v = vb * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)vg * blockDim.x) {
    // every input of the point first: the stores below may alias the
    // inputs as far as the compiler knows, so loads written after a store
    // would wait for it, one memory latency per group
    double rho[2], chi[2], tau[2], lap[2], divj[2], grad[6], cur[6], curls[6], curlj[6];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      rho[q] = a.rho[q * n + v];
      chi[q] = a.chi[q * n + v];
      tau[q] = a.tau[q * n + v];
      lap[q] = a.lap[q * n + v];
      divj[q] = a.divj[q * n + v];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      grad[k] = a.grad[k * n + v];
      cur[k] = a.cur[k * n + v];
      curls[k] = a.curls[k * n + v];
      curlj[k] = a.curlj[k * n + v];
    }
    const double wc = f.tcoul ? a.wcoul[v] : 0.0;
    const double rn = rho[0], rp = rho[1];
    const double rt = rn + rp;
    const double rtp = pow(rt, f.power);
    double rearr = 0.0;
    if (f.ipair == 6) {
      const double cn = chi[0], cp = chi[1];
      rearr = (f.v0neut / f.rho0pr) * cn * cn + (f.v0prot / f.rho0pr) * cp * cp;
      ep[0] += rt * rearr;
    }
#pragma unroll
    for (int iq = 0; iq < 2; ++iq) {
      const int ic = 1 - iq;
      const double rq = rho[iq], rc = rho[ic];
      // meanfield.py:181-190: three-body term
      double u = rtp * ((f.b3 * (f.power + 2.0) / 3.0 - 2.0 * f.b3p / 3.0) * rq + f.b3 * (f.power + 2.0) / 3.0 * rc -
                        (f.b3p * f.power / 3.0) * (rn * rn + rp * rp) / (rt + 1.0e-25));
      // meanfield.py:202-206: spin-orbit divergence
      u += -(f.b4 + f.b4p) * divj[iq] - f.b4 * divj[ic];
      // meanfield.py:211-221: Coulomb and Slater exchange for protons
      if (iq == 1 && f.tcoul) {
        u += wc;
        if (f.ex != 0) u += -f.slate * pow(rp, 1.0 / 3.0);
      }
      // meanfield.py:234-241
      u += (f.b0 - f.b0p) * rq + f.b0 * rc + (f.b1 - f.b1p) * tau[iq] + f.b1 * tau[ic] - (f.b2 - f.b2p) * lap[iq] -
           f.b2 * lap[ic];
      if (f.ipair == 6) u += rearr;  // meanfield.py:352-360
      const double b = f.h2m[iq] + (f.b1 - f.b1p) * rq + f.b1 * rc;  // meanfield.py:253-256
      const double v0 = iq == 1 ? f.v0prot : f.v0neut;
      const double vp = f.ipair == 6 ? v0 * chi[iq] * (1.0 - rt / f.rho0pr) : v0 * chi[iq];
      a.upot[iq * n + v] = u;
      a.bmass[iq * n + v] = b;
      a.vpair[iq * n + v] = vp;
      vm[iq] = fmax(vm[iq], vp);
      if (a.upot32) {
        a.upot32[iq * n + v] = float(u);
        a.bmass32[iq * n + v] = float(b);
        a.vpair32[iq * n + v] = float(vp);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int kq = iq * 3 + c, kc = ic * 3 + c;
        const int64_t iqc = (int64_t)kq * n + v;
        const double w = (f.b4 + f.b4p) * grad[kq] + f.b4 * grad[kc];                      // 270-274
        const double aq = -2.0 * (f.b1 - f.b1p) * cur[kq] - 2.0 * f.b1 * cur[kc] - (f.b4 + f.b4p) * curls[kq] -
                          f.b4 * curls[kc];                                                      // 294-299
        const double sp = -(f.b4 + f.b4p) * curlj[kq] - f.b4 * curlj[kc];                   // 320-324
        a.wlspot[iqc] = w;
        a.aq[iqc] = aq;
        a.spot[iqc] = sp;
        if (a.upot32) {
          a.wlspot32[iqc] = float(w);
          a.aq32[iqc] = float(aq);
          a.spot32[iqc] = float(sp);
        }
      }
    }
  }
  block_sum<1>(ep, red);
  if (threadIdx.x == 0) a.epart[vb] = ep[0];
  // per-block maxima of v_pair
  for (int o = 32; o > 0; o >>= 1) {
    vm[0] = fmax(vm[0], __shfl_xor(vm[0], o, 64));
    vm[1] = fmax(vm[1], __shfl_xor(vm[1], o, 64));
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    vred[wid][0] = vm[0];
    vred[wid][1] = vm[1];
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double m = vred[0][threadIdx.x];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, vred[w][threadIdx.x]);
    a.epart[(1 + threadIdx.x) * vg + vb] = m;
  }
}

__global__ void __launch_bounds__(256) skyrme_combine_kernel(SkyArgs a) { skyrme_body(a, blockIdx.x, gridDim.x); }

// ---------------------------------------------------------------------------
// Energy integrals (energies.py:90-306): per-block partial sums.
// ---------------------------------------------------------------------------
// (workgroup vb of vg)
__device__ __forceinline__ void energy_body(const EnArgs& a, int vb, int vg) {
  const hfb_force& f = a.f;
  const int64_t n = a.nvox;
  __shared__ double red[4 * NES];
  double s[NES];
#pragma unroll
  for (int j = 0; j < NES; ++j) s[j] = 0.0;
  const double sc = f.ex != 0 ? -0.75 * f.slate : 0.0;
  for (int64_t v = vb * (int64_t)blockDim.x + threadIdx.x; v < n; v += (int64_t)vg * blockDim.x) {
    const double rn = a.rho[v], rp = a.rho[n + v];
    const double rt = rn + rp, r1 = -rn + rp;
    const double ln = a.lap[v], lp = a.lap[n + v];
    const double lt = ln + lp, l1 = -ln + lp;
    const double tn = a.tau[v], tp = a.tau[n + v];
    const double t0 = tn + tp, t1 = -tn + tp;
    const double rpow = pow(rt, f.power);
    s[S_EHF0] += (f.b0 * rt * rt - f.b0p * (rp * rp + rn * rn)) / 2.0;
    s[S_CRHO0] += (f.Crho0 + f.Crho0D * rpow) * rt * rt;
    s[S_CRHO1] += (f.Crho1 + f.Crho1D * rpow) * r1 * r1;
    s[S_EHF3] += rpow * (f.b3 * rt * rt - f.b3p * (rp * rp + rn * rn)) / 3.0;
    s[S_EHF2] += (-f.b2 * rt * lt + f.b2p * (rp * lp + rn * ln)) / 2.0;
    s[S_CDRHO0] += f.Cdrho0 * rt * lt;
    s[S_CDRHO1] += f.Cdrho1 * r1 * l1;
    s[S_CTAU0] += f.Ctau0 * t0 * rt;
    s[S_CTAU1] += f.Ctau1 * t1 * r1;
    s[S_EHF1] += f.b1 * rt * t0 - f.b1p * (rn * tn + rp * tp);
    const double dn = a.divj[v], dp = a.divj[n + v];
    s[S_EHFLS] += -f.b4 * rt * (dn + dp) - f.b4p * (rp * dp + rn * dn);
    s[S_CDJ0] += f.CdJ0 * rt * (dn + dp);
    s[S_CDJ1] += f.CdJ1 * r1 * (-dn + dp);
    double odd = 0.0;
    for (int c = 0; c < 3; ++c) {
      const double c0 = a.curlj[(0 * 3 + c) * n + v], c1 = a.curlj[(1 * 3 + c) * n + v];
      const double s0 = a.sdens[(0 * 3 + c) * n + v], s1 = a.sdens[(1 * 3 + c) * n + v];
      odd += -f.b4 * (s0 * (c0 + c1) + s1 * (c0 + c1)) - f.b4p * (s0 * c0 + s1 * c1);
    }
    s[S_LSODD] += odd;
    if (f.tcoul) {
      const double r43 = pow(rp, 4.0 / 3.0);
      s[S_EHFC] += 0.5 * rp * a.wcoul[v] + sc * r43;
      s[S_ECORC] += sc / 3.0 * r43;
    }
    s[S_EHFT] += f.h2m[0] * tn + f.h2m[1] * tp;
  }
  block_sum<NES>(s, red);
  if (threadIdx.x == 0)
    for (int j = 0; j < NES; ++j) a.part[(int64_t)vb * NES + j] = s[j];
}

__global__ void __launch_bounds__(256) energy_kernel(EnArgs a) { energy_body(a, blockIdx.x, gridDim.x); }

// skyrme_combine and the energy partials in one launch (both read only the
// densities, the derived fields and wcoul): workgroups [0, nb) and [nb, 2 nb)
__global__ void __launch_bounds__(256) skyrme_energy_kernel(SkyArgs sa, EnArgs ea, int nb) {
  if ((int)blockIdx.x < nb)
    skyrme_body(sa, blockIdx.x, nb);
  else
    energy_body(ea, blockIdx.x - nb, nb);
}

__global__ void energy_finish_kernel(EnArgs a, int nblk, double wxyz, const double* __restrict__ cur, int nx,
                                     int ny, int nz, double* __restrict__ E) {
  energy_finish_body(a, nblk, wxyz, cur, nx, ny, nz, E);
}

__global__ void ecorrp_finish_kernel(const double* part, int nblk, double wxyz, int ipair, double* E) {
  ecorrp_finish_body(part, nblk, wxyz, ipair, E);
}

__global__ void sum_energy_kernel(int ns, const double* wocc, const double* wstates, const double* kin,
                                  const double* spe, const double* orb, const double* spin, double* E) {
  sum_energy_body(ns, wocc, wstates, kin, spe, orb, spin, E);
}

// The scalar tail of an SCF iteration in one workgroup: ecorrp and the
// v_pair maxima from the skyrme partials, the energy integrals from the
// energy partials, then sum_energy on the single-particle properties.
__global__ void __launch_bounds__(256) energy_tail_kernel(TailArgs t) {
  double vm[2];
  energy_tail_body<4>(t, vm);
}

void push(DerivJobs& J, double* dst, int nterm, const double* s0, int a0, int o0, double c0, const double* s1 = nullptr,
          int a1 = 0, int o1 = 1, double c1 = 0.0, const double* s2 = nullptr, int a2 = 0, int o2 = 1, double c2 = 0.0) {
  DerivJob& j = J.j[J.njob++];
  j.dst = dst;
  j.nterm = nterm;
  j.src[0] = s0; j.axis[0] = a0; j.order[0] = o0; j.coef[0] = c0;
  j.src[1] = s1; j.axis[1] = a1; j.order[1] = o1; j.coef[1] = c1;
  j.src[2] = s2; j.axis[2] = a2; j.order[2] = o2; j.coef[2] = c2;
}

// Workspace slots used by a job list (terms of its multi-term jobs).
int deriv_slots(const DerivJobs& J) {
  int n = 0;
  for (int j = 0; j < J.njob; ++j)
    if (J.j[j].nterm > 1) n += J.j[j].nterm;
  return n;
}

// The term tables of a job list for deriv_tile_body: returns the number of
// terms; *nslot receives the workspace slots of the multi-term jobs.
int prepare_terms(DerivJobs& J, int* nslot) {
  int nt = 0, slot = 0;
  for (int j = 0; j < J.njob; ++j) {
    const bool multi = J.j[j].nterm > 1;
    J.slot0[j] = multi ? slot : -1;
    for (int t = 0; t < J.j[j].nterm; ++t) {
      J.tjob[nt] = j;
      J.tidx[nt] = t;
      J.tslot[nt] = multi ? slot++ : -1;
      ++nt;
    }
  }
  J.nterm_all = nt;
  *nslot = slot;
  return nt;
}

int run_jobs(const hfb_ctx* c, DerivJobs& J, double* work, hipStream_t st) {
  if (c->nx == c->ny && c->ny == c->nz) {
    if (!work && deriv_slots(J) > 0) return HFB_E_ARG;
    DerivMats dm{{c->der1[0], c->der1[1], c->der1[2]}, {c->der2[0], c->der2[1], c->der2[2]}};
    int slot = 0;
    const int nt = prepare_terms(J, &slot);
    const int n = c->nx;
    const int64_t nvox = (int64_t)n * n * n;
    HFB_DISPATCH_N(n, {
      hipLaunchKernelGGL(deriv_tile_kernel<N_>, dim3(N_ * N_ / DT_P, nt), dim3(256), 0, st, J, dm, work);
    })
    int rc = launch_status();
    if (rc != HFB_OK || slot == 0) return rc;
    int gx = (int)((nvox + 255) / 256);
    if (gx > 256) gx = 256;
    hipLaunchKernelGGL(deriv_sum_kernel, dim3(gx, J.njob), dim3(256), 0, st, J, (const double*)work, nvox);
    return launch_status();
  }
  const int64_t nvox = (int64_t)c->nx * c->ny * c->nz;
  const int gx = (int)((nvox + 255) / 256);
  hipLaunchKernelGGL(deriv_kernel, dim3(gx, J.njob), dim3(256), 0, st, J, c->nx, c->ny, c->nz, c->der1[0], c->der1[1],
                     c->der1[2], c->der2[0], c->der2[1], c->der2[2]);
  return launch_status();
}

}  // namespace

// Derived density fields used by skyrme and integ_energy, each (2, ...):
//   divj (2,n3) = div sodens   (meanfield.py:195-200, energies.py:170-176)
//   lap  (2,n3) = Laplacian rho (meanfield.py:227-231, energies.py:95-100)
//   grad (2,3,n3) = grad rho   (meanfield.py:262-268)
//   curls(2,3,n3) = curl sdens (meanfield.py:279-289)
//   curlj(2,3,n3) = curl current (meanfield.py:304-314, energies.py:196-209)
extern "C" long long hfb_deriv_workspace(hfb_ctx* c) {
  if (!c) return HFB_E_NOCTX;
  return 36LL * c->nx * c->ny * c->nz;  // the 36 terms of the multi-term jobs of hfb_density_derivs
}

namespace {

// The derived-field jobs of hfb_density_derivs.
void density_jobs(const hfb_ctx* c, const double* rho, const double* sdens, const double* current,
                  const double* sodens, double* divj, double* lap, double* grad, double* curls, double* curlj,
                  DerivJobs& J) {
  const int64_t n = (int64_t)c->nx * c->ny * c->nz;
  J = DerivJobs{};
  for (int iq = 0; iq < 2; ++iq) {
    const double* r = rho + iq * n;
    const double* so = sodens + iq * 3 * n;
    const double* sd = sdens + iq * 3 * n;
    const double* cu = current + iq * 3 * n;
    push(J, divj + iq * n, 3, so, 0, 1, 1.0, so + n, 1, 1, 1.0, so + 2 * n, 2, 1, 1.0);
    push(J, lap + iq * n, 3, r, 0, 2, 1.0, r, 1, 2, 1.0, r, 2, 2, 1.0);
    for (int a = 0; a < 3; ++a) push(J, grad + (iq * 3 + a) * n, 1, r, a, 1, 1.0);
    // curl v = (dy vz - dz vy, dz vx - dx vz, dx vy - dy vx)
    push(J, curls + (iq * 3 + 0) * n, 2, sd + 2 * n, 1, 1, 1.0, sd + 1 * n, 2, 1, -1.0);
    push(J, curls + (iq * 3 + 1) * n, 2, sd + 0 * n, 2, 1, 1.0, sd + 2 * n, 0, 1, -1.0);
    push(J, curls + (iq * 3 + 2) * n, 2, sd + 1 * n, 0, 1, 1.0, sd + 0 * n, 1, 1, -1.0);
    push(J, curlj + (iq * 3 + 0) * n, 2, cu + 2 * n, 1, 1, 1.0, cu + 1 * n, 2, 1, -1.0);
    push(J, curlj + (iq * 3 + 1) * n, 2, cu + 0 * n, 2, 1, 1.0, cu + 2 * n, 0, 1, -1.0);
    push(J, curlj + (iq * 3 + 2) * n, 2, cu + 1 * n, 0, 1, 1.0, cu + 0 * n, 1, 1, -1.0);
  }
}

// Density derivatives and the Coulomb solve in four launches (cubic grid
// of edge N, doubled grid NP = 2N): the derivative tiles, the forward
// Poisson planes, then the sums of the multi-term fields with the x-pencil
// kernel product (workgroups [0, nbp)), then the inverse planes.  (The
// tiles inside the plane launch ran at one 512-thread workgroup per CU,
// limited by the plane kernel's registers and LDS: slower than apart.)
template <int NP>
__global__ void __launch_bounds__(NP * PP_L) derivs_pois_fwd_kernel(cx<double>* __restrict__ pbuf,
                                                                     const double* __restrict__ rho_p, int nx, int ny,
                                                                     int nz, const cx<double>* __restrict__ tw2) {
  extern __shared__ unsigned char dp_lds[];
  pois_fwd_body<NP>(pbuf, rho_p, nx, ny, nz, tw2, blockIdx.x, (cx<double>*)dp_lds);
}

template <int NP>
__global__ void __launch_bounds__(256) derivs_pois_mul_kernel(DerivJobs jobs, const double* __restrict__ work,
                                                              int64_t nvox, int gx, int nbp,
                                                              cx<double>* __restrict__ pbuf,
                                                              const cx<double>* __restrict__ qhat, double scale,
                                                              int nx, const cx<double>* __restrict__ tw2) {
  if ((int)blockIdx.x < nbp) {
    const int lane = threadIdx.x & 63;
    const int g = ((int)blockIdx.x * 4 + (threadIdx.x >> 6)) * PP_PW + lane / PP_L;
    pois_mul_body<NP>(pbuf, qhat, scale, nx, tw2, g, lane % PP_L);
    return;
  }
  const int b = (int)blockIdx.x - nbp;
  deriv_sum_body(jobs, work, nvox, b % gx, gx, b / gx);
}

template <int NP>
__global__ void __launch_bounds__(NP * PP_L) derivs_pois_inv_kernel(const cx<double>* __restrict__ pbuf,
                                                                     double* __restrict__ wcoul, double scale, int ny,
                                                                     int nz, const cx<double>* __restrict__ tw2) {
  extern __shared__ unsigned char dp_lds[];
  pois_inv_body<NP>(pbuf, wcoul, scale, ny, nz, tw2, blockIdx.x, (cx<double>*)dp_lds);
}

template <int N, int NP>
int derivs_coulomb_fused(const hfb_ctx* c, DerivJobs& J, double* work, const double* rho_p, double* wcoul, double e2,
                         cx<double>* pbuf, hipStream_t st) {
  const cx<double>* tw2 = twiddles<double>(c, NP);
  if (!tw2) return HFB_E_SIZE;
  DerivMats dm{{c->der1[0], c->der1[1], c->der1[2]}, {c->der2[0], c->der2[1], c->der2[2]}};
  int nslot = 0;
  const int nt = prepare_terms(J, &nslot);
  hipLaunchKernelGGL(deriv_tile_kernel<N>, dim3(N * N / DT_P, nt), dim3(256), 0, st, J, dm, work);
  int rc = launch_status();
  if (rc != HFB_OK) return rc;
  const size_t lds = sizeof(cx<double>) * (size_t)NP * (NP + 1);
  allow_full_lds((const void*)derivs_pois_fwd_kernel<NP>);
  allow_full_lds((const void*)derivs_pois_inv_kernel<NP>);
  hipLaunchKernelGGL(derivs_pois_fwd_kernel<NP>, dim3(c->nx), dim3(NP * PP_L), lds, st, pbuf, rho_p, c->nx, c->ny,
                     c->nz, tw2);
  rc = launch_status();
  if (rc != HFB_OK) return rc;
  const int64_t nvox = (int64_t)N * N * N;
  int gx = (int)((nvox + 255) / 256);
  if (gx > 256) gx = 256;
  const int nbp = NP * NP / (4 * PP_PW);
  const int nsum = nslot > 0 ? gx * J.njob : 0;
  hipLaunchKernelGGL(derivs_pois_mul_kernel<NP>, dim3(nbp + nsum), dim3(256), 0, st, J, (const double*)work, nvox, gx,
                     nbp, pbuf, c->qhat, e2 * c->dx * c->dy * c->dz, c->nx, tw2);
  rc = launch_status();
  if (rc != HFB_OK) return rc;
  hipLaunchKernelGGL(derivs_pois_inv_kernel<NP>, dim3(c->nx), dim3(NP * PP_L), lds, st, (const cx<double>*)pbuf, wcoul,
                     1.0 / ((double)NP * NP * NP), c->ny, c->nz, tw2);
  return launch_status();
}

}  // namespace

extern "C" int hfb_density_derivs(hfb_ctx* c, const double* rho, const double* sdens, const double* current,
                                  const double* sodens, double* divj, double* lap, double* grad, double* curls,
                                  double* curlj, double* work, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!rho || !sdens || !current || !sodens || !divj || !lap || !grad || !curls || !curlj) return HFB_E_ARG;
  DerivJobs J;
  density_jobs(c, rho, sdens, current, sodens, divj, lap, grad, curls, curlj, J);
  return run_jobs(c, J, work, (hipStream_t)stream);
}

extern "C" int hfb_density_derivs_coulomb(hfb_ctx* c, const double* rho, const double* sdens, const double* current,
                                          const double* sodens, double* divj, double* lap, double* grad,
                                          double* curls, double* curlj, double* work, double e2, double* wcoul,
                                          void* pwork, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!rho || !sdens || !current || !sodens || !divj || !lap || !grad || !curls || !curlj || !work || !wcoul ||
      !pwork)
    return HFB_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  DerivJobs J;
  density_jobs(c, rho, sdens, current, sodens, divj, lap, grad, curls, curlj, J);
  const bool cubic = c->nx == c->ny && c->ny == c->nz;
  if (cubic && pois_planes_apply(c) && c->nx2 == 2 * c->nx) {
    const int64_t nvox = (int64_t)c->nx * c->ny * c->nz;
    if (c->nx == 32) return derivs_coulomb_fused<32, 64>(c, J, work, rho + nvox, wcoul, e2, (cx<double>*)pwork, st);
    if (c->nx == 48) return derivs_coulomb_fused<48, 96>(c, J, work, rho + nvox, wcoul, e2, (cx<double>*)pwork, st);
  }
  const int rc = run_jobs(c, J, work, st);
  if (rc != HFB_OK) return rc;
  return hfb_poisson(c, rho, wcoul, e2, pwork, stream);
}

// divaq = div aq, dbmass = grad bmass (meanfield.py:329-344); computed for
// completeness of the Meanfield state (hpsi01 does not read them).
extern "C" int hfb_meanfield_derivs(hfb_ctx* c, const double* aq, const double* bmass, double* divaq, double* dbmass,
                                    double* work, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!aq || !bmass || !divaq || !dbmass) return HFB_E_ARG;
  const int64_t n = (int64_t)c->nx * c->ny * c->nz;
  DerivJobs J{};
  for (int iq = 0; iq < 2; ++iq) {
    const double* A = aq + iq * 3 * n;
    push(J, divaq + iq * n, 3, A, 0, 1, 1.0, A + n, 1, 1, 1.0, A + 2 * n, 2, 1, 1.0);
    for (int a = 0; a < 3; ++a) push(J, dbmass + (iq * 3 + a) * n, 1, bmass + iq * n, a, 1, 1.0);
  }
  return run_jobs(c, J, work, (hipStream_t)stream);
}

// Potentials of meanfield.py:163-391 from densities + derived fields +
// wcoul.  The *32 outputs (float copies for the FP32 spinor kernels) may be
// null.  E receives meanfield.ecorrp in slot HFB_EN_ECORRP.  epart: scratch of
// hfb_skyrme_nblocks() doubles.
extern "C" int hfb_skyrme_nblocks(hfb_ctx* c) {
  if (!c) return HFB_E_NOCTX;
  const int64_t n = (int64_t)c->nx * c->ny * c->nz;
  int nb = (int)((n + 255) / 256);
  return nb > 1024 ? 1024 : nb;
}

extern "C" int hfb_skyrme(hfb_ctx* c, const hfb_force* f, const double* rho, const double* tau, const double* chi,
                          const double* current, const double* sdens, const double* divj, const double* lap,
                          const double* grad, const double* curls, const double* curlj, const double* wcoul,
                          double* upot, double* bmass, double* wlspot, double* aq, double* spot, double* v_pair,
                          float* upot32, float* bmass32, float* wlspot32, float* aq32, float* spot32, float* v_pair32,
                          double* epart, double* E, int finish, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!f || !rho || !tau || !chi || !current || !sdens || !divj || !lap || !grad || !curls || !curlj || !upot ||
      !bmass || !wlspot || !aq || !spot || !v_pair || !epart || !E)
    return HFB_E_ARG;
  if (f->tcoul && !wcoul) return HFB_E_ARG;
  if (upot32 && (!bmass32 || !wlspot32 || !aq32 || !spot32 || !v_pair32)) return HFB_E_ARG;
  SkyArgs a{};
  a.f = *f;
  a.nvox = (int64_t)c->nx * c->ny * c->nz;
  a.rho = rho;
  a.tau = tau;
  a.chi = chi;
  a.cur = current;
  a.sdens = sdens;
  a.divj = divj;
  a.lap = lap;
  a.grad = grad;
  a.curls = curls;
  a.curlj = curlj;
  a.wcoul = wcoul;
  a.upot = upot;
  a.bmass = bmass;
  a.wlspot = wlspot;
  a.aq = aq;
  a.spot = spot;
  a.vpair = v_pair;
  a.upot32 = upot32;
  a.bmass32 = bmass32;
  a.wlspot32 = wlspot32;
  a.aq32 = aq32;
  a.spot32 = spot32;
  a.vpair32 = v_pair32;
  a.epart = epart;
  const int nb = hfb_skyrme_nblocks(c);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(skyrme_combine_kernel, dim3(nb), dim3(256), 0, st, a);
  int rc = launch_status();
  if (rc != HFB_OK || !finish) return rc;
  hipLaunchKernelGGL(ecorrp_finish_kernel, dim3(1), dim3(256), 0, st, epart, nb, c->dx * c->dy * c->dz, f->ipair, E);
  return launch_status();
}

// integ_energy: writes the functional integrals into E.  epart: scratch of
// hfb_skyrme_nblocks()*20 doubles.  Pairing energies are read from
// E[HFB_EN_EPAIR..+1].
extern "C" int hfb_integ_energy(hfb_ctx* c, const hfb_force* f, const double* rho, const double* tau,
                                const double* current, const double* sdens, const double* divj, const double* lap,
                                const double* curlj, const double* wcoul, double* epart, double* E, int finish,
                                void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!f || !rho || !tau || !current || !sdens || !divj || !lap || !curlj || !epart || !E) return HFB_E_ARG;
  if (f->tcoul && !wcoul) return HFB_E_ARG;
  EnArgs a{};
  a.f = *f;
  a.nvox = (int64_t)c->nx * c->ny * c->nz;
  a.rho = rho;
  a.tau = tau;
  a.sdens = sdens;
  a.divj = divj;
  a.lap = lap;
  a.curlj = curlj;
  a.wcoul = wcoul;
  a.part = epart;
  const int nb = hfb_skyrme_nblocks(c);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(energy_kernel, dim3(nb), dim3(256), 0, st, a);
  int rc = launch_status();
  if (rc != HFB_OK || !finish) return rc;
  hipLaunchKernelGGL(energy_finish_kernel, dim3(1), dim3(256), 0, st, a, nb, c->dx * c->dy * c->dz, current, c->nx,
                     c->ny, c->nz, E);
  return launch_status();
}

extern "C" int hfb_skyrme_energy(hfb_ctx* c, const hfb_force* f, const double* rho, const double* tau,
                                 const double* chi, const double* current, const double* sdens, const double* divj,
                                 const double* lap, const double* grad, const double* curls, const double* curlj,
                                 const double* wcoul, double* upot, double* bmass, double* wlspot, double* aq,
                                 double* spot, double* v_pair, float* upot32, float* bmass32, float* wlspot32,
                                 float* aq32, float* spot32, float* v_pair32, double* sky_part, double* en_part,
                                 void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!f || !rho || !tau || !chi || !current || !sdens || !divj || !lap || !grad || !curls || !curlj || !upot ||
      !bmass || !wlspot || !aq || !spot || !v_pair || !sky_part || !en_part)
    return HFB_E_ARG;
  if (f->tcoul && !wcoul) return HFB_E_ARG;
  if (upot32 && (!bmass32 || !wlspot32 || !aq32 || !spot32 || !v_pair32)) return HFB_E_ARG;
  SkyArgs a{};
  a.f = *f;
  a.nvox = (int64_t)c->nx * c->ny * c->nz;
  a.rho = rho;
  a.tau = tau;
  a.chi = chi;
  a.cur = current;
  a.sdens = sdens;
  a.divj = divj;
  a.lap = lap;
  a.grad = grad;
  a.curls = curls;
  a.curlj = curlj;
  a.wcoul = wcoul;
  a.upot = upot;
  a.bmass = bmass;
  a.wlspot = wlspot;
  a.aq = aq;
  a.spot = spot;
  a.vpair = v_pair;
  a.upot32 = upot32;
  a.bmass32 = bmass32;
  a.wlspot32 = wlspot32;
  a.aq32 = aq32;
  a.spot32 = spot32;
  a.vpair32 = v_pair32;
  a.epart = sky_part;
  EnArgs e{};
  e.f = *f;
  e.nvox = a.nvox;
  e.rho = rho;
  e.tau = tau;
  e.sdens = sdens;
  e.divj = divj;
  e.lap = lap;
  e.curlj = curlj;
  e.wcoul = wcoul;
  e.part = en_part;
  const int nb = hfb_skyrme_nblocks(c);
  hipLaunchKernelGGL(skyrme_energy_kernel, dim3(2 * nb), dim3(256), 0, (hipStream_t)stream, a, e, nb);
  return launch_status();
}

extern "C" int hfb_energy_tail(hfb_ctx* c, const hfb_force* f, const double* sky_part, const double* en_part,
                               const double* current, int nstates, const double* wocc, const double* wstates,
                               const double* sp_kinetic, const double* sp_energy, const double* sp_orbital,
                               const double* sp_spin, double* E, void* stream) {
  if (!c) return HFB_E_NOCTX;
  const hfb_tail_job j{f, sky_part, en_part, current, nstates, wocc, wstates, sp_kinetic, sp_energy, sp_orbital,
                       sp_spin, E};
  TailArgs t;
  if (!tail_args(c, j, t)) return HFB_E_ARG;
  hipLaunchKernelGGL(energy_tail_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, t);
  return launch_status();
}

extern "C" int hfb_sum_energy(int nstates, const double* wocc, const double* wstates, const double* sp_kinetic,
                              const double* sp_energy, const double* sp_orbital, const double* sp_spin, double* E,
                              void* stream) {
  if (nstates <= 0 || !wocc || !wstates || !sp_kinetic || !sp_energy || !sp_orbital || !sp_spin || !E)
    return HFB_E_ARG;
  hipLaunchKernelGGL(sum_energy_kernel, dim3(1), dim3(256), 0, (hipStream_t)stream, nstates, wocc, wstates,
                     sp_kinetic, sp_energy, sp_orbital, sp_spin, E);
  return launch_status();
}
