// ============================================================================
// reduce.hpp — FP64 device reductions of the 3D solver (forces, Trefftz,
// residual) and helpers shared by the plain and batched translation units.
// Deterministic: fixed block shape and fixed summation order, no atomics.
// ============================================================================
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aero3d {

constexpr int kRedBlock = 256;
constexpr int kRedMaxBlocks = 1024;

inline int reduce_blocks(int n) { int nb = (n + kRedBlock - 1) / kRedBlock; return nb < 1 ? 1 : (nb > kRedMaxBlocks ? kRedMaxBlocks : nb); }

// Hardware dispatches workgroup b to XCD b % 8; give each XCD a contiguous
// range of logical tiles so x-y neighbours (which read each other's halos)
// share an L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int per = nb / 8, rem = nb % 8, x = b % 8, l = b / 8;
  return x < rem ? x * (per + 1) + l : rem * (per + 1) + (x - rem) * per + l;
}

namespace {

// Wavefront (64-lane) then block sum of K doubles; lane 0 of wave 0 holds the
// block total. Fixed shape, fixed order: the result depends only on inputs.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double (*lds)[K]) {
#pragma unroll
  for (int c = 0; c < K; c++)
    for (int off = 32; off > 0; off >>= 1) v[c] += __shfl_down(v[c], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int c = 0; c < K; c++) lds[wave][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < K; c++) {
      double a = 0;
      for (int w = 0; w < kRedBlock / 64; w++) a += lds[w][c];
      v[c] = a;
    }
}

// Pressure force over the force list (cfdForces, src/air_cfd.js:443-449), FP64.
__global__ void __launch_bounds__(kRedBlock) forces_kernel(const float *__restrict__ Q, size_t ns, const int32_t *__restrict__ fq,
                                                           const double *__restrict__ cx, const double *__restrict__ cz,
                                                           const double *__restrict__ x, const double *__restrict__ z, int n,
                                                           double xcg, double zcg, double *__restrict__ partials) {
  __shared__ double lds[kRedBlock / 64][3];
  double v[3] = {0, 0, 0};
  const double pinf = 1.0 / 1.4;
  for (int s = blockIdx.x * kRedBlock + threadIdx.x; s < n; s += gridDim.x * kRedBlock) {
    const int q = fq[s];
    const double r = Q[q], u = Q[ns + q] / r, vv = Q[2 * ns + q] / r, w = Q[3 * ns + q] / r;
    const double dp = (1.4 - 1) * (Q[4 * ns + q] - 0.5 * r * (u * u + vv * vv + w * w)) - pinf;
    const double fx = -dp * cx[s], fz = -dp * cz[s];
    v[0] += fx; v[1] += fz; v[2] += (z[s] - zcg) * fx - (x[s] - xcg) * fz;
  }
  block_sum<3>(v, lds);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; c++) partials[blockIdx.x * 3 + c] = v[c];
}

// Trefftz plane sum (src/air_cfd.js:563-568), FP64, one thread per (j, k).
__global__ void __launch_bounds__(kRedBlock) trefftz_kernel(const float *__restrict__ Q, size_t ns, const uint8_t *__restrict__ t,
                                                            const float *__restrict__ dy, const float *__restrict__ dz,
                                                            int NX, int NY, int NZ, int i0, double tx, double ca, double sa,
                                                            double *__restrict__ partials) {
  __shared__ double lds[kRedBlock / 64][1];
  double v[1] = {0};
  for (int jk = blockIdx.x * kRedBlock + threadIdx.x; jk < NY * NZ; jk += gridDim.x * kRedBlock) {
    const int j = jk % NY, k = jk / NY;
    double u = 0, vv = 0, w = 0, ws = 0;
    for (int di = 0; di < 2; di++) {
      const int q = (k * NY + j) * NX + i0 + di;
      if (t[q] != 0) continue;
      const double ww = di ? tx : 1 - tx, r = Q[q];
      u += ww * Q[ns + q] / r; vv += ww * Q[2 * ns + q] / r; w += ww * Q[3 * ns + q] / r; ws += ww;
    }
    if (ws < 1e-6) continue;
    u /= ws; vv /= ws; w /= ws;
    const double vn = w * ca - u * sa;
    v[0] += (vv * vv + vn * vn) * (double)dy[j] * (double)dz[k];
  }
  block_sum<1>(v, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
}

// Density residual terms (src/air_cfd.js:554): sum of squared change and count
// over every third fluid cell, FP64.
__global__ void __launch_bounds__(kRedBlock) residual_kernel(const float *__restrict__ r, const float *__restrict__ prev,
                                                             const uint8_t *__restrict__ t, int N, double *__restrict__ partials) {
  __shared__ double lds[kRedBlock / 64][2];
  double v[2] = {0, 0};
  const int M = (N + 2) / 3;
  for (int m = blockIdx.x * kRedBlock + threadIdx.x; m < M; m += gridDim.x * kRedBlock) {
    const int q = 3 * m;
    if (t[q] != 0) continue;
    const double d = (double)r[q] - (double)prev[q];
    v[0] += d * d; v[1] += 1;
  }
  block_sum<2>(v, lds);
  if (threadIdx.x == 0) { partials[blockIdx.x * 2] = v[0]; partials[blockIdx.x * 2 + 1] = v[1]; }
}

// Final fixed-order sum of nb block partials of K components.
template <int K>
__global__ void __launch_bounds__(kRedBlock) finish_kernel(const double *__restrict__ partials, int nb, double *__restrict__ out) {
  __shared__ double lds[kRedBlock / 64][K];
  double v[K];
#pragma unroll
  for (int c = 0; c < K; c++) v[c] = 0;
  for (int b = threadIdx.x; b < nb; b += kRedBlock)
#pragma unroll
    for (int c = 0; c < K; c++) v[c] += partials[b * K + c];
  block_sum<K>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < K; c++) out[c] = v[c];
}


}  // namespace
}  // namespace aero3d
