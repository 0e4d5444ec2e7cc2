// ============================================================================
// reduce.hpp — FP64 device reductions of the 3D solver: the readback sums
// (density residual, force list, Trefftz plane, Mach maximum) and helpers
// shared by the plain and batched translation units.
//
// Deterministic: fixed block shape, fixed block count per sum (a function of
// the sum's length only) and fixed summation order, no atomics. The readback
// of one solve is a segment; a launch covers any number of segments, and a
// segment's partial and final sums depend only on that segment, so one solve
// read back alone (plain path) or among many (served form) gives the same
// bits. Per-term arithmetic is that of the reference's double loops with no
// multiply-add contraction, so only the order of the grid-wide sums differs
// from the reference's sequential loops.
// ============================================================================
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
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

// Readback of one solve (CFD3.readback, src/air_cfd.js:552-558): the density
// residual since the snapshot and the force-list sums, then the coefficients.
struct RbSeg {
  const float *Q;               // current state, 5 planes of ns
  unsigned long long ns;
  const uint8_t *t;
  int N;
  float *rsnap;                 // density snapshot; every third cell is read and rewritten
  int nit;                      // iterations since the snapshot; 0: take the snapshot only (res = NaN)
  const int32_t *fq;            // force list (cfdForceList, src/air_cfd.js:143-152)
  const double *cx, *cz, *x, *z;
  int nf;
  double xcg, zcg;
  double ca, sa, kq, cref;      // cos, sin of alpha; 2 / (M^2 S / 2); reference chord
  int nb_res, nb_f;             // blocks of the residual and force sums (reduce_blocks of their lengths)
  int blk0;                     // first block of this segment in a segmented launch
  int slot;                     // served form: the case's status record (unused by the plain path)
  int it;                       // iterations of the case at this readback
  int rule, cap;                // served form: convergence rule and iteration cap
  double res_drop;
  const int *skip;              // non-null and set: the segment is skipped (a finished served case)
  double *part;                 // partials: residual 2 nb_res, then forces 3 nb_f
  double *out;                  // kRbOut results
};
// out[]: sum of squared density change, count, Fx, Fz, My, res, CL, CD, Cm
constexpr int kRbOut = 9;

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
  __syncthreads();   // lds may be reused by the caller's next sum
}

// Block b of nb of the density residual (src/air_cfd.js:554): squared change
// and count over every third fluid cell; the snapshot is rewritten with the
// current density at every third cell (the only cells the residual reads).
__device__ __forceinline__ void residual_block(const RbSeg &g, int b, int nb, double *part, double (*lds)[2]) {
#pragma clang fp contract(off)
  double v[2] = {0, 0};
  const int M = (g.N + 2) / 3;
  for (int m = b * kRedBlock + (int)threadIdx.x; m < M; m += nb * kRedBlock) {
    const int q = 3 * m;
    const float r = g.Q[q];
    if (g.nit > 0 && g.t[q] == 0) {
      const double d = (double)r - (double)g.rsnap[q];
      v[0] += d * d; v[1] += 1;
    }
    g.rsnap[q] = r;
  }
  block_sum<2>(v, lds);
  if (threadIdx.x == 0) { part[2 * b] = v[0]; part[2 * b + 1] = v[1]; }
}

// Block b of nb of the pressure force over the force list (cfdForces,
// src/air_cfd.js:443-449).
__device__ __forceinline__ void forces_block(const RbSeg &g, int b, int nb, double *part, double (*lds)[3]) {
#pragma clang fp contract(off)
  double v[3] = {0, 0, 0};
  const double pinf = 1.0 / 1.4;
  const size_t ns = g.ns;
  for (int s = b * kRedBlock + (int)threadIdx.x; s < g.nf; s += nb * kRedBlock) {
    const int q = g.fq[s];
    const double r = g.Q[q], u = g.Q[ns + q] / r, vv = g.Q[2 * ns + q] / r, w = g.Q[3 * ns + q] / r;
    const double dp = (1.4 - 1) * (g.Q[4 * ns + q] - 0.5 * r * (u * u + vv * vv + w * w)) - pinf;
    const double fx = -dp * g.cx[s], fz = -dp * g.cz[s];
    v[0] += fx; v[1] += fz; v[2] += (g.z[s] - g.zcg) * fx - (g.x[s] - g.xcg) * fz;
  }
  block_sum<3>(v, lds);
  if (threadIdx.x == 0)
    for (int c = 0; c < 3; c++) part[3 * b + c] = v[c];
}

// Fixed-order sum of nb block partials of K components into out[0..K).
template <int K>
__device__ __forceinline__ void finish_sum(const double *part, int nb, double *out, double (*lds)[K]) {
  double v[K];
#pragma unroll
  for (int c = 0; c < K; c++) v[c] = 0;
  for (int b = threadIdx.x; b < nb; b += kRedBlock)
#pragma unroll
    for (int c = 0; c < K; c++) v[c] += part[b * K + c];
  block_sum<K>(v, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int c = 0; c < K; c++) out[c] = v[c];
}

// Segment of block b of a segmented launch (segments in block order).
__device__ __forceinline__ int seg_of_block(const RbSeg *seg, int nseg, int b) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) { const int m = (lo + hi + 1) >> 1; if (seg[m].blk0 <= b) lo = m; else hi = m - 1; }
  return __builtin_amdgcn_readfirstlane(lo);
}

// Partial sums of every segment: blocks [blk0, blk0 + nb_res) the residual,
// then nb_f blocks of the force list.
__global__ void __launch_bounds__(kRedBlock) rb_partials_kernel(const RbSeg *__restrict__ seg, int nseg) {
  __shared__ double l2[kRedBlock / 64][2];
  __shared__ double l3[kRedBlock / 64][3];
  const int s = seg_of_block(seg, nseg, blockIdx.x);
  const RbSeg g = seg[s];
  if (g.skip && __builtin_amdgcn_readfirstlane(*g.skip)) return;
  const int lb = blockIdx.x - g.blk0;
  if (lb < g.nb_res) residual_block(g, lb, g.nb_res, g.part, l2);
  else forces_block(g, lb - g.nb_res, g.nb_f, g.part + 2 * g.nb_res, l3);
}

// Residual and coefficients of a segment from its sums (thread 0).
__device__ __forceinline__ void rb_results(const RbSeg &g) {
#pragma clang fp contract(off)
  double *o = g.out;
  o[5] = g.nit > 0 ? sqrt(o[0] / fmax(1.0, o[1])) / g.nit : (double)NAN;
  // wind-axis rotation and half-model doubling (src/air_cfd.js:450-451)
  o[6] = g.kq * (o[3] * g.ca - o[2] * g.sa);
  o[7] = g.kq * (o[2] * g.ca + o[3] * g.sa);
  o[8] = g.kq * o[4] / g.cref;
}

// One block per segment: final sums and results.
__global__ void __launch_bounds__(kRedBlock) rb_finish_kernel(const RbSeg *__restrict__ seg) {
  __shared__ double l2[kRedBlock / 64][2];
  __shared__ double l3[kRedBlock / 64][3];
  const RbSeg g = seg[blockIdx.x];
  finish_sum<2>(g.part, g.nb_res, g.out, l2);
  finish_sum<3>(g.part + 2 * g.nb_res, g.nb_f, g.out + 2, l3);
  if (threadIdx.x == 0) rb_results(g);
}

// Trefftz plane sum (src/air_cfd.js:563-568), FP64, one thread per (j, k).
__global__ void __launch_bounds__(kRedBlock) trefftz_kernel(const float *__restrict__ Q, size_t ns, const uint8_t *__restrict__ t,
                                                            const float *__restrict__ dy, const float *__restrict__ dz,
                                                            int NX, int NY, int NZ, int i0, double tx, double ca, double sa,
                                                            double *__restrict__ partials) {
#pragma clang fp contract(off)
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

// Final fixed-order sum of nb block partials of K components.
template <int K>
__global__ void __launch_bounds__(kRedBlock) finish_kernel(const double *__restrict__ partials, int nb, double *__restrict__ out) {
  __shared__ double lds[kRedBlock / 64][K];
  finish_sum<K>(partials, nb, out, lds);
}

// Squared Mach number maximum of the readback (src/air_cfd.js:571): over every
// second fluid cell, (u^2 + v^2 + w^2) / (1.4 p / r) in doubles; a larger value
// replaces the running one only when it compares greater (a NaN never does),
// so the result is independent of the order and equals the sequential loop's.
__device__ __forceinline__ double gtmax(double a, double b) { return b > a ? b : a; }
__global__ void __launch_bounds__(kRedBlock) machmax_kernel(const float *__restrict__ Q, size_t ns, const uint8_t *__restrict__ t, int N,
                                                            double *__restrict__ partials) {
#pragma clang fp contract(off)
  __shared__ double lds[kRedBlock / 64];
  double m = 0;
  const int M = (N + 1) / 2;
  for (int h = blockIdx.x * kRedBlock + threadIdx.x; h < M; h += gridDim.x * kRedBlock) {
    const int q = 2 * h;
    if (t[q] != 0) continue;
    const double r = Q[q], u = Q[ns + q] / r, v = Q[2 * ns + q] / r, w = Q[3 * ns + q] / r;
    const double p = (1.4 - 1) * (Q[4 * ns + q] - 0.5 * r * (u * u + v * v + w * w));
    const double m2 = (u * u + v * v + w * w) / (1.4 * p / r);
    m = gtmax(m, m2);
  }
  for (int off = 32; off > 0; off >>= 1) m = gtmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = lds[0];
    for (int w = 1; w < kRedBlock / 64; w++) a = gtmax(a, lds[w]);
    partials[blockIdx.x] = a;
  }
}
__global__ void machmax_finish(const double *__restrict__ partials, int nb, double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double a = 0;
  for (int b = 0; b < nb; b++) a = gtmax(a, partials[b]);
  *out = a;
}

}  // namespace
}  // namespace aero3d
