// Densities summed over states (densities.py:177-337) fused with the
// per-state single-particle properties (inout.py:64-149).
//
// One launch per axis a.  A workgroup owns one pencil tile of one isospin
// and walks the states of that isospin, NSUB states at a time (one state
// per sub-group of NTS threads).  For each state the tile of both spin
// components is transformed along a in LDS; the spectrum gives the state's
// kinetic-energy contribution sum k^2 |psi_k|^2 / n (Nyquist kept, the
// Parseval form of inout.py:86-91,130) and the inverse transform of
// i k psi_k gives d_a psi (Nyquist zeroed, cderv*00/01's d1).  Density
// contributions accumulate in registers of the points each thread owns;
// sub-groups are combined in LDS at the end, so no per-state field is ever
// written to HBM.  The x pass also adds the pointwise rho, chi, spin
// density and the per-state spin and parity sums.
//
// Density partials (double) are written per axis into slots of
// `dout` = [2 iso][17 slot][nvox]:
//   0 rho 1 chi 2 sx 3 sy 4 sz 5 tau_x 6 j_x 7 Jy(x) 8 Jz(x)     (x pass)
//   9 tau_y 10 j_y 11 Jx(y) 12 Jz(y)                              (y pass)
//   13 tau_z 14 j_z 15 Jx(z) 16 Jy(z)                             (z pass)
// and combined by hfb_density_finish (with the linear mixing of
// static.py:1817-1823 when requested).  Per-state sums go to
// spart = [3 axis][S][ntile_a][8]: kin, Lx, Ly, Lz, sx, sy, sz, parity.
#include <math.h>

#include "hfb_ctx.h"

using namespace hfb;

namespace {

constexpr int NSLOT = 17;

struct DensArgs {
  int nx, ny, nz;
  double dx, dy, dz;
  const void* psi;
  int nst0, nst1;  // states per isospin
  const double* wocc;
  const double* wstates;
  const double* wguv;
  const double* pairwg;
  double cm[3];
  double* dout;   // [2][17][nvox]
  double* spart;  // [S][ntile][8] for this axis
  int want_dens;
  int want_sp;
  int ntile;
};

__device__ __forceinline__ double im_ab(double ax, double ay, double bx, double by) {  // Im(a conj b)
  return ay * bx - ax * by;
}
__device__ __forceinline__ double re_ab(double ax, double ay, double bx, double by) {  // Re(a conj b)
  return ax * bx + ay * by;
}

template <typename T, int N, int AXIS, int P, int NSUB>
__global__ void __launch_bounds__(NSUB * 8 * P) dens_axis_kernel(DensArgs a, const cx<T>* __restrict__ twg) {
  constexpr int Q = 2 * P;
  constexpr int NTS = 8 * P;          // threads per sub-group
  constexpr int NT = NSUB * NTS;
  constexpr int PPT = (N * P + NTS - 1) / NTS;  // tile points per thread
  constexpr int NACC = AXIS == 0 ? 9 : 4;
  extern __shared__ unsigned char lds_raw[];
  __shared__ cx<T> tw[N];
  __shared__ double red[NT / 64 * 8];
  cx<T>* base = (cx<T>*)lds_raw;
  const int sg = threadIdx.x / NTS;
  const int lt = threadIdx.x % NTS;
  cx<T>* sPSI = base + (size_t)sg * 3 * N * Q;
  cx<T>* sW0 = sPSI + N * Q;
  cx<T>* sW1 = sW0 + N * Q;

  const int tile = blockIdx.x;
  const int iq = blockIdx.y;
  const int nst = iq == 0 ? a.nst0 : a.nst1;
  const int start = iq == 0 ? 0 : a.nst0;
  const int64_t nvox = (int64_t)a.nx * a.ny * a.nz;
  const Tile t = make_tile<AXIS>(tile, a.nx, a.ny, a.nz, P);
  const double dax = AXIS == 0 ? a.dx : (AXIS == 1 ? a.dy : a.dz);
  const T kfac = T(2.0 * M_PI / (dax * N));
  const double kf = 2.0 * M_PI / (dax * N);
  const cx<T>* psi = (const cx<T>*)a.psi;
  load_twiddles(tw, twg, N);

  double acc[PPT][NACC];
#pragma unroll
  for (int i = 0; i < PPT; ++i)
#pragma unroll
    for (int j = 0; j < NACC; ++j) acc[i][j] = 0.0;

  for (int s0 = 0; s0 < nst; s0 += NSUB) {
    const int sl = s0 + sg;
    const bool active = sl < nst;
    const int s = start + (active ? sl : 0);
    const cx<T>* p0 = psi + (int64_t)(2 * s) * nvox;
    const cx<T>* p1 = p0 + nvox;
    for (int e = lt; e < N * P; e += NTS) {
      int k, p;
      elem_of<AXIS, N, P>(e, k, p);
      const int64_t off = t.base + k * t.sk + p * t.sp;
      cx<T> u = {T(0), T(0)}, v = {T(0), T(0)};
      if (active) {
        u = p0[off];
        v = p1[off];
      }
      sPSI[k * Q + p] = u;
      sPSI[k * Q + P + p] = v;
      sW0[k * Q + p] = u;
      sW0[k * Q + P + p] = v;
    }
    __syncthreads();
    cx<T>* r = fft_lds<T, N, Q, -1>(sW0, sW1, tw, Q, lt, NTS);
    double sp[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = lt; e < N * Q; e += NTS) {
      const int k = e / Q;
      const cx<T> v = r[e];
      const double m = double(k <= N / 2 ? k : N - k) * kf;
      sp[0] += m * m * (double(v.x) * double(v.x) + double(v.y) * double(v.y));
      r[e] = cmul_is(v, d1_factor<T, N>(k, kfac));
    }
    sp[0] *= 1.0 / N;
    __syncthreads();
    cx<T>* d1 = fft_lds<T, N, Q, +1>(r, r == sW0 ? sW1 : sW0, tw, Q, lt, NTS);
    const double w = active ? a.wocc[s] * a.wstates[s] : 0.0;
    const double wuv = active ? a.wguv[s] * a.pairwg[s] * a.wstates[s] : 0.0;
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int e = lt + i * NTS;
      if (e >= N * P) break;
      int k, p;
      elem_of<AXIS, N, P>(e, k, p);
      const cx<T> u0 = sPSI[k * Q + p], u1 = sPSI[k * Q + P + p];
      const cx<T> g0 = d1[k * Q + p], g1 = d1[k * Q + P + p];
      const double ax = u0.x, ay = u0.y, bx = u1.x, by = u1.y;
      const double cx0 = g0.x, cy0 = g0.y, cx1 = g1.x, cy1 = g1.y;
      const double tau = cx0 * cx0 + cy0 * cy0 + cx1 * cx1 + cy1 * cy1;
      const double ia = im_ab(cx0, cy0, ax, ay) + im_ab(cx1, cy1, bx, by);  // Im(d psi conj psi)
      int x, y, z;
      coords_of<AXIS>(tile, k, p, a.nx, a.ny, a.nz, P, x, y, z);
      const double xc = (x - 0.5 * (a.nx - 1)) * a.dx - a.cm[0];
      const double yc = (y - 0.5 * (a.ny - 1)) * a.dy - a.cm[1];
      const double zc = (z - 0.5 * (a.nz - 1)) * a.dz - a.cm[2];
      acc[i][0] += w * tau;
      acc[i][1] += w * ia;
      if constexpr (AXIS == 0) {
        // densities.py:270-277
        acc[i][2] += -w * (im_ab(cx0, cy0, ax, ay) - im_ab(cx1, cy1, bx, by));   // J_y
        acc[i][3] += -w * (re_ab(ax, ay, cx1, cy1) - re_ab(bx, by, cx0, cy0));   // J_z
        const double n0 = ax * ax + ay * ay, n1 = bx * bx + by * by;
        const double rsx = ax * bx + ay * by;  // Re(conj psi0 psi1)
        const double rsy = ax * by - ay * bx;  // Im(conj psi0 psi1)
        acc[i][4] += w * (n0 + n1);
        acc[i][5] += 0.5 * wuv * (n0 + n1);
        acc[i][6] += 2.0 * w * rsx;
        acc[i][7] += 2.0 * w * rsy;
        acc[i][8] += w * (n0 - n1);
        sp[2] += zc * ia;
        sp[3] -= yc * ia;
        sp[4] += 2.0 * rsx;
        sp[5] += 2.0 * rsy;
        sp[6] += n0 - n1;
        if (active) {  // parity: Re sum conj psi(r) psi(-r)  (inout.py:134-137)
          const int64_t mo = ((int64_t)(a.nx - 1 - x) * a.ny + (a.ny - 1 - y)) * a.nz + (a.nz - 1 - z);
          const cx<T> f0 = p0[mo], f1 = p1[mo];
          sp[7] += ax * double(f0.x) + ay * double(f0.y) + bx * double(f1.x) + by * double(f1.y);
        }
      } else if constexpr (AXIS == 1) {
        // densities.py:289-296
        acc[i][2] += w * (im_ab(cx0, cy0, ax, ay) - im_ab(cx1, cy1, bx, by));    // J_x
        acc[i][3] += -w * (im_ab(cx1, cy1, ax, ay) + im_ab(cx0, cy0, bx, by));   // J_z
        sp[1] -= zc * ia;
        sp[3] += xc * ia;
      } else {
        // densities.py:308-315
        acc[i][2] += w * (re_ab(cx1, cy1, ax, ay) - re_ab(cx0, cy0, bx, by));    // J_x
        acc[i][3] += w * (im_ab(cx1, cy1, ax, ay) + im_ab(cx0, cy0, bx, by));    // J_y
        sp[1] += yc * ia;
        sp[2] -= xc * ia;
      }
    }
    if (a.want_sp) {
      // sub-group sums (each sub-group = NTS/64 waves)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        double v = sp[j];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if ((threadIdx.x & 63) == 0) red[(threadIdx.x >> 6) * 8 + j] = v;
      }
      __syncthreads();
      if (lt < 8 && active) {
        double v = 0.0;
        for (int wv = 0; wv < NTS / 64; ++wv) v += red[(sg * (NTS / 64) + wv) * 8 + lt];
        a.spart[((int64_t)s * a.ntile + tile) * 8 + lt] = v;
      }
    }
    __syncthreads();
  }
  if (!a.want_dens) return;
  // combine the sub-groups' accumulators through LDS, slot by slot
  double* comb = (double*)lds_raw;  // NSUB * N*P doubles (fits in the tile buffers)
  const int slot0 = AXIS == 0 ? 0 : (AXIS == 1 ? 9 : 13);
  // accumulator j -> slot: x: tau,j,Jy,Jz,rho,chi,sx,sy,sz -> 5,6,7,8,0,1,2,3,4
  for (int j = 0; j < NACC; ++j) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PPT; ++i) {
      const int e = lt + i * NTS;
      if (e < N * P) comb[sg * N * P + e] = acc[i][j];
    }
    __syncthreads();
    int slot;
    if constexpr (AXIS == 0)
      slot = j < 4 ? 5 + j : j - 4;
    else
      slot = slot0 + j;
    double* dst = a.dout + ((int64_t)iq * NSLOT + slot) * nvox;
    for (int e = threadIdx.x; e < N * P; e += NT) {
      double v = 0.0;
#pragma unroll
      for (int g = 0; g < NSUB; ++g) v += comb[g * N * P + e];
      int k, p;
      elem_of<AXIS, N, P>(e, k, p);
      dst[t.base + k * t.sk + p * t.sp] = v;
    }
  }
}

template <typename T, int AXIS, int P, int NSUB>
int launch_dens(const hfb_ctx* c, DensArgs a, hipStream_t st) {
  const int n = axis_len(c, AXIS);
  const int nt = tiles_per_field<AXIS>(c->nx, c->ny, c->nz, P);
  a.ntile = nt;
  HFB_DISPATCH_N(n, {
    constexpr int NN = N_;
    const size_t lds = sizeof(cx<T>) * (size_t)NSUB * 3 * NN * 2 * P;
    auto kern = dens_axis_kernel<T, NN, AXIS, P, NSUB>;
    if (lds > 32 * 1024) allow_full_lds((const void*)kern);
    hipLaunchKernelGGL(kern, dim3(nt, 2), dim3(NSUB * 8 * P), lds, st, a, twiddles<T>(c, NN));
  })
  return launch_status();
}

// Sum the axis slots, optional mixing new*addnew + old*(1-addnew) of rho,
// tau, chi (static.py:1817-1823).
__global__ void dens_finish_kernel(const double* __restrict__ d, int64_t nvox, double* rho, double* chi,
                                   double* tau, double* cur, double* sdens, double* sodens, int mix, double addnew) {
  const int64_t tot = 2 * nvox;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot; e += (int64_t)gridDim.x * blockDim.x) {
    const int iq = e / nvox;
    const int64_t v = e % nvox;
    const double* s = d + (int64_t)iq * NSLOT * nvox + v;
    double r = s[0 * nvox], ch = s[1 * nvox];
    double ta = s[5 * nvox] + s[9 * nvox] + s[13 * nvox];
    if (mix) {
      const double ac = 1.0 - addnew;
      r = addnew * r + ac * rho[e];
      ta = addnew * ta + ac * tau[e];
      ch = addnew * ch + ac * chi[e];
    }
    rho[e] = r;
    chi[e] = ch;
    tau[e] = ta;
    const int64_t b3 = (int64_t)iq * 3 * nvox + v;
    cur[b3] = s[6 * nvox];
    cur[b3 + nvox] = s[10 * nvox];
    cur[b3 + 2 * nvox] = s[14 * nvox];
    sdens[b3] = s[2 * nvox];
    sdens[b3 + nvox] = s[3 * nvox];
    sdens[b3 + 2 * nvox] = s[4 * nvox];
    sodens[b3] = s[11 * nvox] + s[15 * nvox];
    sodens[b3 + nvox] = s[7 * nvox] + s[16 * nvox];
    sodens[b3 + 2 * nvox] = s[8 * nvox] + s[12 * nvox];
  }
}

// Per-state properties from the axis partials (inout.py:98-148).
__global__ void sp_finish_kernel(const double* __restrict__ spart, int ns, int nst0, int nt0, int nt1, int nt2,
                                 double wxyz, double h2m0, double h2m1, double* kin, double* orb, double* spin,
                                 double* par) {
  const int s = blockIdx.x;
  __shared__ double red[4 * 8];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int nts[3] = {nt0, nt1, nt2};
  int64_t off = 0;
  for (int ax = 0; ax < 3; ++ax) {
    for (int tt = threadIdx.x; tt < nts[ax]; tt += blockDim.x) {
      const double* p = spart + off + ((int64_t)s * nts[ax] + tt) * 8;
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] += p[j];
    }
    off += (int64_t)ns * nts[ax] * 8;
  }
  block_sum<8>(v, red);
  if (threadIdx.x == 0) {
    const double h2m = s < nst0 ? h2m0 : h2m1;
    kin[s] = wxyz * h2m * v[0];
    orb[3 * s + 0] = wxyz * v[1];
    orb[3 * s + 1] = wxyz * v[2];
    orb[3 * s + 2] = wxyz * v[3];
    spin[3 * s + 0] = wxyz * 0.5 * v[4];
    spin[3 * s + 1] = wxyz * 0.5 * v[5];
    spin[3 * s + 2] = wxyz * 0.5 * v[6];
    par[s] = wxyz * v[7];
  }
}

}  // namespace

extern "C" int hfb_density_sizes(hfb_ctx* c, int prec, long long* dout_doubles, long long* spart_doubles_per_state) {
  if (!c) return HFB_E_NOCTX;
  const int P = 8;
  const long long nvox = (long long)c->nx * c->ny * c->nz;
  if (dout_doubles) *dout_doubles = 2LL * NSLOT * nvox;
  if (spart_doubles_per_state)
    *spart_doubles_per_state = 8LL * (tiles_per_field<0>(c->nx, c->ny, c->nz, P) +
                                      tiles_per_field<1>(c->nx, c->ny, c->nz, P) +
                                      tiles_per_field<2>(c->nx, c->ny, c->nz, P));
  (void)prec;
  return HFB_OK;
}

// want_dens: write the density slots to dout; want_sp: per-state sums.
extern "C" int hfb_density(hfb_ctx* c, int prec, const void* psi, int nst0, int nst1, const double* wocc,
                           const double* wstates, const double* wguv, const double* pairwg, const double* cmtot,
                           int want_dens, int want_sp, double* dout, double* spart, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!psi || !wocc || !wstates || !wguv || !pairwg || nst0 < 0 || nst1 < 0) return HFB_E_ARG;
  if ((want_dens && !dout) || (want_sp && !spart)) return HFB_E_ARG;
  constexpr int P = 8, NSUB = 4;
  DensArgs a{};
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.dx = c->dx;
  a.dy = c->dy;
  a.dz = c->dz;
  a.psi = psi;
  a.nst0 = nst0;
  a.nst1 = nst1;
  a.wocc = wocc;
  a.wstates = wstates;
  a.wguv = wguv;
  a.pairwg = pairwg;
  a.cm[0] = cmtot ? cmtot[0] : 0.0;
  a.cm[1] = cmtot ? cmtot[1] : 0.0;
  a.cm[2] = cmtot ? cmtot[2] : 0.0;
  a.dout = dout;
  a.want_dens = want_dens;
  a.want_sp = want_sp;
  const int ns = nst0 + nst1;
  const int nt0 = tiles_per_field<0>(c->nx, c->ny, c->nz, P);
  const int nt1 = tiles_per_field<1>(c->nx, c->ny, c->nz, P);
  hipStream_t st = (hipStream_t)stream;
  int rc = HFB_OK;
  for (int ax = 0; ax < 3 && rc == HFB_OK; ++ax) {
    a.spart = spart ? spart + (ax == 0 ? 0 : (ax == 1 ? (int64_t)ns * nt0 * 8 : (int64_t)ns * (nt0 + nt1) * 8)) : nullptr;
    if (prec == HFB_F64) {
      if (ax == 0) rc = launch_dens<double, 0, P, NSUB>(c, a, st);
      if (ax == 1) rc = launch_dens<double, 1, P, NSUB>(c, a, st);
      if (ax == 2) rc = launch_dens<double, 2, P, NSUB>(c, a, st);
    } else if (prec == HFB_F32) {
      if (ax == 0) rc = launch_dens<float, 0, P, NSUB>(c, a, st);
      if (ax == 1) rc = launch_dens<float, 1, P, NSUB>(c, a, st);
      if (ax == 2) rc = launch_dens<float, 2, P, NSUB>(c, a, st);
    } else {
      return HFB_E_ARG;
    }
  }
  return rc;
}

extern "C" int hfb_density_finish(hfb_ctx* c, const double* dout, double* rho, double* chi, double* tau,
                                  double* current, double* sdens, double* sodens, int mix, double addnew,
                                  void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!dout || !rho || !chi || !tau || !current || !sdens || !sodens) return HFB_E_ARG;
  const int64_t nvox = (int64_t)c->nx * c->ny * c->nz;
  const int nb = (int)((2 * nvox + 255) / 256);
  hipLaunchKernelGGL(dens_finish_kernel, dim3(nb), dim3(256), 0, (hipStream_t)stream, dout, nvox, rho, chi, tau,
                     current, sdens, sodens, mix, addnew);
  return launch_status();
}

extern "C" int hfb_sp_finish(hfb_ctx* c, const double* spart, int nst0, int nst1, double h2m0, double h2m1,
                             double* sp_kinetic, double* sp_orbital, double* sp_spin, double* sp_parity,
                             void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!spart || !sp_kinetic || !sp_orbital || !sp_spin || !sp_parity) return HFB_E_ARG;
  constexpr int P = 8;
  const int ns = nst0 + nst1;
  if (ns <= 0) return HFB_E_ARG;
  hipLaunchKernelGGL(sp_finish_kernel, dim3(ns), dim3(256), 0, (hipStream_t)stream, spart, ns, nst0,
                     tiles_per_field<0>(c->nx, c->ny, c->nz, P), tiles_per_field<1>(c->nx, c->ny, c->nz, P),
                     tiles_per_field<2>(c->nx, c->ny, c->nz, P), c->dx * c->dy * c->dz, h2m0, h2m1, sp_kinetic,
                     sp_orbital, sp_spin, sp_parity);
  return launch_status();
}
