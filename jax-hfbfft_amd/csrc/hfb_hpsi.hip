// One-axis spectral derivatives and the fused single-particle Hamiltonian.
//
// h psi (meanfield.py:40-159) is built in three pencil passes, one per axis.
// Each pass stages a tile of pencils of both spin components in LDS, runs
// the FFT derivatives of that axis there (cderv02's d1 and d(B d1) with the
// Nyquist rule, and cderv00 of the spin-orbit/current combination), and adds
// its terms to the accumulator.  The x pass also adds the pointwise U and
// spin-current terms; the z pass finishes with the weighting
// pout = w h psi - wuv v_pair psi and the per-tile overlap partial sums.
#include <math.h>

#include "hfb_ctx.h"

using namespace hfb;

namespace {

template <typename T>
__device__ __forceinline__ cx<T>* other_of(cx<T>* p, cx<T>* a, cx<T>* b) {
  return p == a ? b : a;
}

// ---------------------------------------------------------------------------
// cderv{x,y,z}{00,01,02} on independent fields (levels.py:131-395)
// ---------------------------------------------------------------------------
template <typename T, int N, int AXIS, int P, int KIND>
__global__ void __launch_bounds__(128) cderv_kernel(const cx<T>* __restrict__ in, const T* __restrict__ B,
                                                    cx<T>* __restrict__ o1, cx<T>* __restrict__ o2, int nx,
                                                    int ny, int nz, T d, const cx<T>* __restrict__ twg) {
  __shared__ cx<T> tw[N];
  __shared__ cx<T> buf[4][N * P];
  __shared__ T mb[P];
  __shared__ cx<T> nyq[P];
  const int tile = blockIdx.x;
  const int64_t nvox = (int64_t)nx * ny * nz;
  const int64_t foff = (int64_t)blockIdx.y * nvox;
  const Tile t = make_tile<AXIS>(tile, nx, ny, nz, P);
  load_twiddles(tw, twg, N);
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    int k, p;
    elem_of<AXIS, N, P>(e, k, p);
    buf[0][k * P + p] = in[foff + t.base + k * t.sk + p * t.sp];
  }
  if constexpr (KIND == 2) {
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
      T s = T(0);
      for (int k = 0; k < N; ++k) s += B[t.base + k * t.sk + p * t.sp];
      mb[p] = s / T(N);
    }
  }
  __syncthreads();
  const T kfac = T(2.0 * M_PI) / (d * T(N));
  cx<T>* r = fft_lds<T, N, P, -1>(buf[0], buf[1], tw, P);
  cx<T>* o = other_of(r, buf[0], buf[1]);
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    const int k = e / P;
    const cx<T> v = r[e];
    if constexpr (KIND == 1) buf[2][e] = cscale(v, -k2_factor<T, N>(k, kfac) / T(N));
    if constexpr (KIND == 2) {
      if (k == N / 2) nyq[e % P] = v;
    }
    r[e] = cmul_is(v, d1_factor<T, N>(k, kfac));
  }
  __syncthreads();
  cx<T>* d1 = fft_lds<T, N, P, +1>(r, o, tw, P);
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    int k, p;
    elem_of<AXIS, N, P>(e, k, p);
    const cx<T> v = d1[k * P + p];
    o1[foff + t.base + k * t.sk + p * t.sp] = v;
    if constexpr (KIND == 2) buf[2][k * P + p] = cscale(v, B[t.base + k * t.sk + p * t.sp]);
  }
  __syncthreads();
  if constexpr (KIND == 1) {
    cx<T>* d2 = fft_lds<T, N, P, +1>(buf[2], buf[3], tw, P);
    for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
      int k, p;
      elem_of<AXIS, N, P>(e, k, p);
      o2[foff + t.base + k * t.sk + p * t.sp] = d2[k * P + p];
    }
  }
  if constexpr (KIND == 2) {
    cx<T>* g = fft_lds<T, N, P, -1>(buf[2], buf[3], tw, P);
    const T nq = -T(M_PI / (double)d) * T(M_PI / (double)d) / T(N);
    for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
      const int k = e / P, p = e % P;
      if (k == N / 2)
        g[e] = cscale(nyq[p], nq * mb[p]);
      else
        g[e] = cmul_is(g[e], d1_factor<T, N>(k, kfac));
    }
    __syncthreads();
    cx<T>* d2 = fft_lds<T, N, P, +1>(g, other_of(g, buf[2], buf[3]), tw, P);
    for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
      int k, p;
      elem_of<AXIS, N, P>(e, k, p);
      o2[foff + t.base + k * t.sk + p * t.sp] = d2[k * P + p];
    }
  }
}

template <typename T, int N, int AXIS, int KIND>
int launch_cderv(const hfb_ctx* c, const void* in, const void* B, void* o1, void* o2, int nf,
                 hipStream_t st) {
  constexpr int P = 8;
  const int nt = tiles_per_field<AXIS>(c->nx, c->ny, c->nz, P);
  hipLaunchKernelGGL((cderv_kernel<T, N, AXIS, P, KIND>), dim3(nt, nf), dim3(128), 0, st,
                     (const cx<T>*)in, (const T*)B, (cx<T>*)o1, (cx<T>*)o2, c->nx, c->ny, c->nz,
                     T(axis_d(c, AXIS)), twiddles<T>(c, N));
  return launch_status();
}

template <typename T>
int cderv_dispatch(const hfb_ctx* c, int axis, int kind, const void* in, const void* B, void* o1, void* o2,
                   int nf, hipStream_t st) {
  const int n = axis_len(c, axis);
#define HFB_CD(AX)                                                                        \
  if (axis == AX) {                                                                       \
    HFB_DISPATCH_N(n, {                                                                   \
      if (kind == 0) return launch_cderv<T, N_, AX, 0>(c, in, B, o1, o2, nf, st);         \
      if (kind == 1) return launch_cderv<T, N_, AX, 1>(c, in, B, o1, o2, nf, st);         \
      return launch_cderv<T, N_, AX, 2>(c, in, B, o1, o2, nf, st);                        \
    })                                                                                    \
  }
  HFB_CD(0)
  HFB_CD(1)
  HFB_CD(2)
#undef HFB_CD
  return HFB_E_ARG;
}

// ---------------------------------------------------------------------------
// Fused h psi, one axis per launch.
// ---------------------------------------------------------------------------
template <typename T>
struct MF {
  const T* U;  // (2, nvox)
  const T* B;  // (2, nvox)
  const T* A;  // (2, 3, nvox)
  const T* S;  // (2, 3, nvox)
  const T* W;  // (2, 3, nvox)
  const T* V;  // (2, nvox)
};

template <typename T, int N, int AXIS, int P, int NT>
__global__ void __launch_bounds__(NT) hpsi_axis_kernel(const cx<T>* __restrict__ psi, cx<T>* __restrict__ pmf,
                                                       cx<T>* __restrict__ pout, cx<T>* __restrict__ pdel,
                                                       MF<T> mf, const double* __restrict__ w,
                                                       const double* __restrict__ wuv, int npsi0, int nx, int ny,
                                                       int nz, T d, const cx<T>* __restrict__ twg,
                                                       double* __restrict__ partial) {
  constexpr int Q = 2 * P;  // pencils: spin 0 in q < P, spin 1 in q >= P
  __shared__ cx<T> tw[N];
  __shared__ cx<T> sPSI[N * Q];
  __shared__ cx<T> sD1[N * Q];
  __shared__ cx<T> sW[3][N * Q];
  __shared__ T mb[P];
  __shared__ cx<T> nyq[Q];
  __shared__ double red[(NT / 64) * 2];

  const int tile = blockIdx.x;
  const int s = blockIdx.y;
  const int iq = s >= npsi0 ? 1 : 0;
  const int64_t nvox = (int64_t)nx * ny * nz;
  const cx<T>* ps0 = psi + (int64_t)(2 * s) * nvox;
  const cx<T>* ps1 = ps0 + nvox;
  const Tile t = make_tile<AXIS>(tile, nx, ny, nz, P);
  const T* Bq = mf.B + iq * nvox;
  const T* Aa = mf.A + (iq * 3 + AXIS) * nvox;
  const T* W0 = mf.W + (iq * 3 + 0) * nvox;
  const T* W1 = mf.W + (iq * 3 + 1) * nvox;
  const T* W2 = mf.W + (iq * 3 + 2) * nvox;

  load_twiddles(tw, twg, N);
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    int k, p;
    elem_of<AXIS, N, P>(e, k, p);
    const int64_t off = t.base + k * t.sk + p * t.sp;
    const cx<T> a = ps0[off], b = ps1[off];
    sPSI[k * Q + p] = a;
    sPSI[k * Q + P + p] = b;
    sW[0][k * Q + p] = a;
    sW[0][k * Q + P + p] = b;
  }
  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    T acc = T(0);
    for (int k = 0; k < N; ++k) acc += Bq[t.base + k * t.sk + p * t.sp];
    mb[p] = acc / T(N);
  }
  __syncthreads();

  const T kfac = T(2.0 * M_PI) / (d * T(N));
  // d1 = d/da psi (cderv00 / the first half of cderv02)
  cx<T>* r = fft_lds<T, N, Q, -1>(sW[0], sW[1], tw, Q);
  for (int e = threadIdx.x; e < N * Q; e += blockDim.x) {
    const int k = e / Q;
    const cx<T> v = r[e];
    if (k == N / 2) nyq[e % Q] = v;
    r[e] = cmul_is(v, d1_factor<T, N>(k, kfac));
  }
  __syncthreads();
  cx<T>* d1p = fft_lds<T, N, Q, +1>(r, other_of(r, sW[0], sW[1]), tw, Q);
  // keep d1; stage B*d1
  for (int e = threadIdx.x; e < N * Q; e += blockDim.x) {
    const int k = e / Q, p = (e % Q) % P;
    const cx<T> v = d1p[e];
    sD1[e] = v;
    sW[2][e] = cscale(v, Bq[t.base + k * t.sk + p * t.sp]);
  }
  __syncthreads();
  // spectrum of d(B d1): ik multiplier, m=0 zero, Nyquist rule (levels.py:309-315)
  cx<T>* x2 = fft_lds<T, N, Q, -1>(sW[2], sW[0], tw, Q);
  const T nq = -T(M_PI / (double)d) * T(M_PI / (double)d) / T(N);
  for (int e = threadIdx.x; e < N * Q; e += blockDim.x) {
    const int k = e / Q, q = e % Q;
    if (k == N / 2)
      x2[e] = cscale(nyq[q], nq * mb[q % P]);
    else
      x2[e] = cmul_is(x2[e], d1_factor<T, N>(k, kfac));
  }
  // the spin-orbit / current combination whose derivative is added
  cx<T>* f1 = (x2 == sW[2]) ? sW[0] : sW[2];
  cx<T>* f2 = sW[1];
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    const int k = e / P, p = e % P;
    const int64_t off = t.base + k * t.sk + p * t.sp;
    const cx<T> a = sPSI[k * Q + p], b = sPSI[k * Q + P + p];
    const T av = Aa[off];
    cx<T> q0, q1;
    if constexpr (AXIS == 0) {  // meanfield.py:75-83
      const T w1 = W1[off], w2 = W2[off];
      q0 = csub(cmul_is(a, T(-0.5) * (av - w1)), cscale(b, T(0.5) * w2));
      q1 = cadd(cmul_is(b, T(-0.5) * (av + w1)), cscale(a, T(0.5) * w2));
    } else if constexpr (AXIS == 1) {  // meanfield.py:104-112
      const T w0 = W0[off], w2 = W2[off];
      q0 = cadd(cmul_is(a, T(-0.5) * (av + w0)), cmul_is(b, T(0.5) * w2));
      q1 = cadd(cmul_is(b, T(-0.5) * (av - w0)), cmul_is(a, T(0.5) * w2));
    } else {  // meanfield.py:133-141
      const T w0 = W0[off], w1 = W1[off];
      q0 = cadd(cmul_is(a, T(-0.5) * av), cadd(cscale(b, T(0.5) * w0), cmul_is(b, T(-0.5) * w1)));
      q1 = cadd(cmul_is(b, T(-0.5) * av), cadd(cscale(a, T(-0.5) * w0), cmul_is(a, T(-0.5) * w1)));
    }
    f1[k * Q + p] = q0;
    f1[k * Q + P + p] = q1;
  }
  __syncthreads();
  cx<T>* x3 = fft_lds<T, N, Q, -1>(f1, f2, tw, Q);
  for (int e = threadIdx.x; e < N * Q; e += blockDim.x) {
    const int k = e / Q;
    x3[e] = csub(cmul_is(x3[e], d1_factor<T, N>(k, kfac)), x2[e]);
  }
  __syncthreads();
  cx<T>* dd = fft_lds<T, N, Q, +1>(x3, other_of(x3, f1, f2), tw, Q);  // d3 - d2

  double acc[2] = {0.0, 0.0};
  const int64_t f0 = (int64_t)(2 * s) * nvox;
  for (int e = threadIdx.x; e < N * P; e += blockDim.x) {
    int k, p;
    elem_of<AXIS, N, P>(e, k, p);
    const int64_t off = t.base + k * t.sk + p * t.sp;
    const cx<T> a = sPSI[k * Q + p], b = sPSI[k * Q + P + p];
    const cx<T> da = sD1[k * Q + p], db = sD1[k * Q + P + p];
    cx<T> o0 = dd[k * Q + p], o1 = dd[k * Q + P + p];
    const T av = Aa[off];
    if constexpr (AXIS == 0) {  // meanfield.py:63-73
      const T w1 = W1[off], w2 = W2[off];
      o0 = cadd(o0, csub(cmul_is(da, -(T(0.5) * av - T(0.5) * w1)), cscale(db, T(0.5) * w2)));
      o1 = cadd(o1, cadd(cmul_is(db, -(T(0.5) * av + T(0.5) * w1)), cscale(da, T(0.5) * w2)));
      // meanfield.py:47-58: U psi and the spin-current coupling
      const T u = mf.U[iq * nvox + off];
      const T s0 = mf.S[(iq * 3 + 0) * nvox + off];
      const T s1 = mf.S[(iq * 3 + 1) * nvox + off];
      const T s2 = mf.S[(iq * 3 + 2) * nvox + off];
      o0 = cadd(o0, cadd(cscale(a, u + s2), cadd(cscale(b, s0), cmul_is(b, -s1))));
      o1 = cadd(o1, cadd(cscale(b, u - s2), cadd(cscale(a, s0), cmul_is(a, s1))));
    } else if constexpr (AXIS == 1) {  // meanfield.py:92-102
      const T w0 = W0[off], w2 = W2[off];
      o0 = cadd(o0, cadd(cmul_is(da, -(T(0.5) * av + T(0.5) * w0)), cmul_is(db, T(0.5) * w2)));
      o1 = cadd(o1, cadd(cmul_is(db, -(T(0.5) * av - T(0.5) * w0)), cmul_is(da, T(0.5) * w2)));
    } else {  // meanfield.py:121-131
      const T w0 = W0[off], w1 = W1[off];
      o0 = cadd(o0, cadd(cmul_is(da, -T(0.5) * av), cadd(cscale(db, T(0.5) * w0), cmul_is(db, T(-0.5) * w1))));
      o1 = cadd(o1, cadd(cmul_is(db, -T(0.5) * av), cadd(cscale(da, T(-0.5) * w0), cmul_is(da, T(-0.5) * w1))));
    }
    if constexpr (AXIS != 0) {
      o0 = cadd(o0, pmf[f0 + off]);
      o1 = cadd(o1, pmf[f0 + nvox + off]);
    }
    pmf[f0 + off] = o0;
    pmf[f0 + nvox + off] = o1;
    if constexpr (AXIS == 2) {  // meanfield.py:150-157
      const T v = mf.V[iq * nvox + off];
      const cx<T> e0 = cscale(a, v), e1 = cscale(b, v);
      if (pdel) {
        pdel[f0 + off] = e0;
        pdel[f0 + nvox + off] = e1;
      }
      const T ws = T(w[s]), wu = T(wuv[s]);
      const cx<T> r0 = csub(cscale(o0, ws), cscale(e0, wu));
      const cx<T> r1 = csub(cscale(o1, ws), cscale(e1, wu));
      if (pout) {
        pout[f0 + off] = r0;
        pout[f0 + nvox + off] = r1;
      }
      acc[0] += double(a.x) * double(o0.x) + double(a.y) * double(o0.y) + double(b.x) * double(o1.x) +
                double(b.y) * double(o1.y);
      acc[1] += double(a.x) * double(r0.x) + double(a.y) * double(r0.y) + double(b.x) * double(r1.x) +
                double(b.y) * double(r1.y);
    }
  }
  if constexpr (AXIS == 2) {
    if (partial) {
      block_sum<2>(acc, red);
      if (threadIdx.x == 0) {
        const int ntile = gridDim.x;
        partial[((int64_t)s * ntile + tile) * 2 + 0] = acc[0];
        partial[((int64_t)s * ntile + tile) * 2 + 1] = acc[1];
      }
    }
  }
}

template <typename T>
int pencils_per_tile(const hfb_ctx* c) {
  if (sizeof(T) == 4 && c->nx % 16 == 0 && c->ny % 16 == 0 && c->nz % 16 == 0) return 16;
  return 8;
}

template <typename T, int AXIS, int P>
int launch_hpsi_axis(const hfb_ctx* c, const void* psi, int ns, int npsi0, MF<T> mf, const double* w,
                     const double* wuv, void* pout, void* pmf, void* pdel, double* partial, hipStream_t st) {
  constexpr int NT = 16 * P;  // 2P pencils x 8 radix-4 butterflies of a 32-point pass
  const int n = axis_len(c, AXIS);
  const int nt = tiles_per_field<AXIS>(c->nx, c->ny, c->nz, P);
  HFB_DISPATCH_N(n, {
    hipLaunchKernelGGL((hpsi_axis_kernel<T, N_, AXIS, P, NT>), dim3(nt, ns), dim3(NT), 0, st, (const cx<T>*)psi,
                       (cx<T>*)pmf, (cx<T>*)pout, (cx<T>*)pdel, mf, w, wuv, npsi0, c->nx, c->ny, c->nz,
                       T(axis_d(c, AXIS)), twiddles<T>(c, N_), partial);
  })
  return launch_status();
}

template <typename T, int P>
int hpsi_all(const hfb_ctx* c, const void* psi, int ns, int npsi0, MF<T> mf, const double* w,
             const double* wuv, void* pout, void* pmf, void* pdel, double* partial, hipStream_t st) {
  int rc = launch_hpsi_axis<T, 0, P>(c, psi, ns, npsi0, mf, w, wuv, pout, pmf, pdel, partial, st);
  if (rc == HFB_OK) rc = launch_hpsi_axis<T, 1, P>(c, psi, ns, npsi0, mf, w, wuv, pout, pmf, pdel, partial, st);
  if (rc == HFB_OK) rc = launch_hpsi_axis<T, 2, P>(c, psi, ns, npsi0, mf, w, wuv, pout, pmf, pdel, partial, st);
  return rc;
}

}  // namespace

extern "C" int hfb_cderv(hfb_ctx* c, int prec, int axis, int kind, const void* in, const void* bmass,
                         void* out1, void* out2, int nfields, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!in || !out1 || nfields <= 0 || axis < 0 || axis > 2 || kind < 0 || kind > 2) return HFB_E_ARG;
  if (kind >= 1 && !out2) return HFB_E_ARG;
  if (kind == 2 && !bmass) return HFB_E_ARG;
  if (prec == HFB_F64) return cderv_dispatch<double>(c, axis, kind, in, bmass, out1, out2, nfields, (hipStream_t)stream);
  if (prec == HFB_F32) return cderv_dispatch<float>(c, axis, kind, in, bmass, out1, out2, nfields, (hipStream_t)stream);
  return HFB_E_ARG;
}

extern "C" int hfb_hpsi_ntiles(hfb_ctx* c, int prec) {
  if (!c) return HFB_E_NOCTX;
  const int P = prec == HFB_F32 ? pencils_per_tile<float>(c) : pencils_per_tile<double>(c);
  return tiles_per_field<2>(c->nx, c->ny, c->nz, P);
}

extern "C" int hfb_hpsi(hfb_ctx* c, int prec, const void* psi, int nstates, int npsi0, const void* upot,
                        const void* bmass, const void* aq, const void* spot, const void* wlspot, const void* v_pair,
                        const double* w, const double* wuv, void* pout, void* pout_mf, void* pout_del,
                        double* partial, int* ntile_out, void* stream) {
  if (!c) return HFB_E_NOCTX;
  if (!psi || !pout_mf || !upot || !bmass || !aq || !spot || !wlspot || !v_pair || !w || !wuv || nstates <= 0)
    return HFB_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (ntile_out) *ntile_out = hfb_hpsi_ntiles(c, prec);
  if (prec == HFB_F64) {
    MF<double> mf{(const double*)upot, (const double*)bmass, (const double*)aq,
                  (const double*)spot, (const double*)wlspot, (const double*)v_pair};
    return hpsi_all<double, 8>(c, psi, nstates, npsi0, mf, w, wuv, pout, pout_mf, pout_del, partial, st);
  }
  if (prec == HFB_F32) {
    MF<float> mf{(const float*)upot, (const float*)bmass, (const float*)aq,
                 (const float*)spot, (const float*)wlspot, (const float*)v_pair};
    if (pencils_per_tile<float>(c) == 16)
      return hpsi_all<float, 16>(c, psi, nstates, npsi0, mf, w, wuv, pout, pout_mf, pout_del, partial, st);
    return hpsi_all<float, 8>(c, psi, nstates, npsi0, mf, w, wuv, pout, pout_mf, pout_del, partial, st);
  }
  return HFB_E_ARG;
}
