// The three launches of the Coulomb solve on a cubic doubled grid
// (coulomb.py:54-75) as workgroup bodies, shared by the Poisson launches
// and by the density-derivative launches that carry the first two.
#pragma once
#include "hfb_ctx.h"

namespace hfb {

constexpr int PP_L = 8;  // lanes per pencil
constexpr int PP_PW = 64 / PP_L;  // x pencils per 64-lane wavefront

// Forward y-z plane of x plane x (< nx) by a workgroup of N * PP_L threads:
// z transforms of the rows y < ny (zero-padded beyond nz), rows y >= ny
// zero, then y transforms; out[x][ky][kz].  pl: N (N + 1) complex of LDS.
template <int N>
__device__ __forceinline__ void pois_fwd_body(cx<double>* __restrict__ out, const double* __restrict__ rho, int nx,
                                              int ny, int nz, const cx<double>* __restrict__ tw, int x,
                                              cx<double>* pl) {
  constexpr int L = PP_L, N1 = N / L, P = N + 1;
  const int j = threadIdx.x % L, q = threadIdx.x / L;
  LaneFFT<double, N, L> F;
  F.init(j, tw);
  cx<double> v[N1];
  if (q < ny) {  // z pencil q = y
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int z = j + L * m;
      v[m] = {z < nz ? rho[((int64_t)x * ny + q) * nz + z] : 0.0, 0.0};
    }
    F.fwd(v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) pl[q * P + F.freq(m)] = v[m];
  } else {
#pragma unroll
    for (int m = 0; m < N1; ++m) pl[q * P + j + L * m] = {0.0, 0.0};
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < N1; ++m) v[m] = pl[(j + L * m) * P + q];  // y pencil q = kz
  F.fwd(v, tw);
#pragma unroll
  for (int m = 0; m < N1; ++m) out[((int64_t)x * N + F.freq(m)) * N + q] = v[m];
}

// x pencil g = ky * N + kz by the PP_L lanes j of a group: forward with
// x >= nx zero, times scale * qhat, inverse, rows x < nx written back.
template <int N>
__device__ __forceinline__ void pois_mul_body(cx<double>* __restrict__ buf, const cx<double>* __restrict__ qhat,
                                              double scale, int nx, const cx<double>* __restrict__ tw, int g, int j) {
  constexpr int L = PP_L, N1 = N / L;
  const int64_t plane = (int64_t)N * N;
  LaneFFT<double, N, L> F;
  F.init(j, tw);
  cx<double> v[N1];
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const int x = j + L * m;
    v[m] = x < nx ? buf[x * plane + g] : cx<double>{0.0, 0.0};
  }
  F.fwd(v, tw);
#pragma unroll
  for (int m = 0; m < N1; ++m) v[m] = cmul(v[m], cscale(qhat[F.freq(m) * plane + g], scale));
  F.inv(v, tw);
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const int x = j + L * m;
    if (x < nx) buf[x * plane + g] = v[m];
  }
}

// Inverse y-z plane of x plane x (< nx) by N * PP_L threads: y transforms
// of every kz, rows y < ny kept in LDS, z transforms of those rows,
// wcoul[x][y][z] for z < nz (real part, times scale = 1 / N^3).
template <int N>
__device__ __forceinline__ void pois_inv_body(const cx<double>* __restrict__ in, double* __restrict__ wcoul,
                                              double scale, int ny, int nz, const cx<double>* __restrict__ tw,
                                              int x, cx<double>* pl) {
  constexpr int L = PP_L, N1 = N / L, P = N + 1;
  const int j = threadIdx.x % L, q = threadIdx.x / L;
  LaneFFT<double, N, L> F;
  F.init(j, tw);
  cx<double> v[N1];
#pragma unroll
  for (int m = 0; m < N1; ++m) v[m] = in[((int64_t)x * N + F.freq(m)) * N + q];  // y pencil q = kz
  F.inv(v, tw);
#pragma unroll
  for (int m = 0; m < N1; ++m) {
    const int y = j + L * m;
    if (y < ny) pl[y * P + q] = v[m];
  }
  __syncthreads();
  if (q < ny) {  // z pencil q = y
#pragma unroll
    for (int m = 0; m < N1; ++m) v[m] = pl[q * P + F.freq(m)];
    F.inv(v, tw);
#pragma unroll
    for (int m = 0; m < N1; ++m) {
      const int z = j + L * m;
      if (z < nz) wcoul[((int64_t)x * ny + q) * nz + z] = v[m].x * scale;
    }
  }
}

// The three-launch solve applies: a cubic doubled grid of a supported edge.
inline bool pois_planes_apply(const hfb_ctx* c) {
  return !c->periodic && c->nx2 == c->ny2 && c->ny2 == c->nz2 && (c->nx2 == 48 || c->nx2 == 64 || c->nx2 == 96);
}

}  // namespace hfb
