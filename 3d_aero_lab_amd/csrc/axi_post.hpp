// ============================================================================
// axi_post.hpp — post-processing of the axisymmetric solver on the device,
// in FP64 on the FP32 state:
//   derived_kernel   the engine page's display fields (cfdDerived,
//                    src/cfd_post.js:10-40): region and material codes,
//                    T, Tt, P, Pt, |V|, M, rho, Vx, Vth, Vr, Y, the azimuthal
//                    vorticity and the entropy change against the freestream
//   interp_kernel    the grid-sequencing transfer of a state onto another
//                    grid (interpState, src/app_cfd.js:321-358)
// Each cell's value is the reference's double arithmetic in its order
// (contraction off, JavaScript max / min / clamp); results are rounded to FP32
// once, where the reference stores them in Float32Arrays (display fields) or
// where the solver's state holds them (transfer).
// ============================================================================
#pragma once
#include "axi_kernels.hpp"

namespace axik {

constexpr int kDerivedFields = 13;   // T, Tt, P, Pt, V, M, rho, Vx, Vth, Vr, Y, vort, ds

// Float32 axial and radial velocity of cell k, as the display fields store them
__device__ __forceinline__ void vel_f32(const float *__restrict__ Q, int N, int k, float &u, float &v) {
#pragma clang fp contract(off)
  const double r = jmax((double)Q[k], 1e-4);
  u = (float)((double)Q[N + k] / r);
  v = (float)((double)Q[2 * N + k] / r);
}

// one thread per cell. solid: the solver's mask, or (disp) also cells under
// half full; stream / mat: zone codes; fields: kDerivedFields planes of N.
__global__ void __launch_bounds__(kB) derived_kernel(AxiView v, const double *__restrict__ xcD, const double *__restrict__ rcD,
                                                     const float *__restrict__ Q, const int8_t *__restrict__ stream,
                                                     const uint8_t *__restrict__ mat, int disp, double T0, double P0,
                                                     float *__restrict__ fields, uint8_t *__restrict__ reg, uint8_t *__restrict__ matout) {
#pragma clang fp contract(off)
  const int k = blockIdx.x * kB + threadIdx.x;
  const int N = v.N, W = v.W, H = v.H;
  if (k >= N) return;
  auto solid = [&](int c) { return v.solid[c] != 0 || (disp && v.phi[c] < 0.5f); };
  float *f[kDerivedFields];
  for (int c = 0; c < kDerivedFields; c++) f[c] = fields + (size_t)c * N;
  const float nan = __builtin_nanf("");
  if (solid(k)) {
    reg[k] = 1;
    matout[k] = mat[k];
    for (int c = 0; c < kDerivedFields; c++) f[c][k] = c == 8 ? 0.f : nan;   // Vth stays 0
    return;
  }
  const int st = stream[k];
  reg[k] = st < 0 ? 0 : (uint8_t)(2 + st);
  matout[k] = 0;
  const PrimD q = prim_d(Q, N, k);
  const double T = q.p / (q.r * kRGas), cp = q.g * kRGas / (q.g - 1.0), V2 = q.u * q.u + q.v * q.v, a = sqrt(q.g * q.p / q.r);
  const double Tt = T + V2 / (2.0 * cp);
  f[0][k] = (float)T;
  f[1][k] = (float)Tt;
  f[2][k] = (float)q.p;
  f[3][k] = (float)(q.p * pow(Tt / T, q.g / (q.g - 1.0)));
  f[4][k] = (float)sqrt(V2);
  f[5][k] = (float)(sqrt(V2) / a);
  f[6][k] = (float)q.r;
  f[7][k] = (float)q.u;
  f[8][k] = 0.f;
  f[9][k] = (float)q.v;
  f[10][k] = (float)q.Y;
  f[12][k] = (float)(cp * log(T / T0) - kRGas * log(q.p / P0));
  // vorticity du/dr - dv/dx, central differences over the non-solid neighbours
  // of the FP32 velocities (one-sided at solid cells and the grid edge)
  const int i = k % W, j = k / W;
  const int kl = i > 0 && !solid(k - 1) ? k - 1 : k, kr = i < W - 1 && !solid(k + 1) ? k + 1 : k;
  const int kd = j > 0 && !solid(k - W) ? k - W : k, ku = j < H - 1 && !solid(k + W) ? k + W : k;
  float ul, vl, ur, vr, ud, vd, uu, vu;
  vel_f32(Q, N, kl, ul, vl); vel_f32(Q, N, kr, ur, vr); vel_f32(Q, N, kd, ud, vd); vel_f32(Q, N, ku, uu, vu);
  const double dxx = xcD[kr % W] - xcD[kl % W], drr = rcD[ku / W] - rcD[kd / W];
  const double dvdx = dxx > 0 ? ((double)vr - (double)vl) / dxx : 0.0, dudr = drr > 0 ? ((double)uu - (double)ud) / drr : 0.0;
  f[11][k] = (float)(dudr - dvdx);
}

// fractional index of v in the ascending centres c[0..n), extrapolated
// linearly beyond the first and last centre
__device__ inline double frac_index(const double *__restrict__ c, int n, double v) {
#pragma clang fp contract(off)
  if (v <= c[0]) return (v - c[0]) / (c[1] - c[0]);
  if (v >= c[n - 1]) return n - 1 + (v - c[n - 1]) / (c[n - 1] - c[n - 2]);
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int m = (lo + hi) >> 1;
    if (c[m] <= v) lo = m; else hi = m;
  }
  return lo + (v - c[lo]) / (c[hi] - c[lo]);
}

struct InterpSrc {
  int W, H;
  const double *xc, *rc;
  const uint8_t *solid;
  const float *Q;                 // 5 planes of W H
};

// one thread per target cell (i, j) of a WB x HB grid: bilinear in the source
// centres over non-solid source cells, renormalised; with no weight, the
// nearest non-solid source cell of the 7 x 7 window (index distance); with
// none, density P0 / (R T0) and energy P0 / 0.4 at rest. QB: 5 planes of WB HB.
__global__ void __launch_bounds__(kB) interp_kernel(InterpSrc A, int WB, int HB, const double *__restrict__ xcB,
                                                    const double *__restrict__ rcB, double T0, double P0, float *__restrict__ QB) {
#pragma clang fp contract(off)
  const int kB_ = blockIdx.x * kB + threadIdx.x;
  const int NB = WB * HB;
  if (kB_ >= NB) return;
  const int i = kB_ % WB, j = kB_ / WB, NA = A.W * A.H;
  const double fj = frac_index(A.rc, A.H, rcB[j]), fi = frac_index(A.xc, A.W, xcB[i]);
  auto clampi = [](int x, int lo, int hi) { return x < lo ? lo : x > hi ? hi : x; };
  auto floor_in = [](double x, int hi) { const double f = floor(x); return f > 0 ? (f < hi ? (int)f : hi) : 0; };   // clamp(floor(x), 0, hi)
  const int i0 = floor_in(fi, A.W - 1), j0 = floor_in(fj, A.H - 1);
  const double tx = jclamp(fi - i0, 0.0, 1.0), ty = jclamp(fj - j0, 0.0, 1.0);
  const double wc[4] = {(1 - tx) * (1 - ty), tx * (1 - ty), (1 - tx) * ty, tx * ty};
  double acc[5] = {0, 0, 0, 0, 0}, wsum = 0.0;
  for (int c = 0; c < 4; c++) {
    const int ii = min(i0 + (c & 1), A.W - 1), jj = min(j0 + (c >> 1), A.H - 1), k = jj * A.W + ii;
    if (A.solid[k] || wc[c] <= 0) continue;
    for (int q = 0; q < 5; q++) acc[q] += wc[c] * (double)A.Q[(size_t)q * NA + k];
    wsum += wc[c];
  }
  if (wsum > 0) {
    for (int q = 0; q < 5; q++) QB[(size_t)q * NB + kB_] = (float)(acc[q] / wsum);
    return;
  }
  int best = -1, bd = 1 << 30;
  for (int dj = -3; dj <= 3; dj++)
    for (int di = -3; di <= 3; di++) {
      const int ii = clampi(i0 + di, 0, A.W - 1), jj = clampi(j0 + dj, 0, A.H - 1), k = jj * A.W + ii;
      if (!A.solid[k] && di * di + dj * dj < bd) { bd = di * di + dj * dj; best = k; }
    }
  for (int q = 0; q < 5; q++) {
    float x;
    if (best >= 0) x = A.Q[(size_t)q * NA + best];
    else x = q == 0 ? (float)(P0 / (kRGas * T0)) : q == 3 ? (float)(P0 / 0.4) : 0.f;
    QB[(size_t)q * NB + kB_] = x;
  }
}

}  // namespace axik
