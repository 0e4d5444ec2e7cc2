// ============================================================================
// probe.hpp — display sampling of a 3D solve at scattered points, one thread
// per point (aero3d_probe).
//
// A point (x, y, z) of the full model maps to (x, |y|, z) of the half domain.
// Its value is the trilinear blend of the eight cell centres around it, over
// the cells that carry a state (fluid, t = 0, and fixed-state cells, t >= 3),
// with the weights renormalised to sum 1 (cfdStencil, src/air_cfd.js:454-469).
// When those cells hold (almost) none of the weight the point takes the
// nearest fluid centre of the 4 x 4 x 4 block around the bracket instead.
// The blended conserved state gives the primitives rho, u, v, w, p, with v
// mirrored below the symmetry plane (cfdSample, :471-477), and from those the
// dimensional display state (cfdState, :484-490). The particle velocity
// (cfdVelAt, :637-646) is the same blend inside the box of cell centres with
// no nearest-cell fallback, divided by the freestream Mach number.
//
// Weights are held in FP32 as the reference's callers store them (Float32Array
// stencils); every sum and quotient is FP64 in the reference's order, so
// stencils and primitives equal the reference's bit for bit. The dimensional
// state goes through sqrt and pow and may differ from a JS engine's in the
// last bit.
// ============================================================================
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace aero3d {

struct ProbeArgs {
  int NX, NY, NZ, n;
  const float *xc, *yc, *zc;     // FP32 cell centres of each axis (each axis >= 2 cells)
  const uint8_t *t;              // cell flags
  const float *Q;                // conserved state, 5 planes of ns floats
  size_t ns;
  const double *pts;             // 3 n coordinates
  double fsT, fsa, fsP, fsrho, fsq, M;   // freestream record of the display state, Mach number
  int32_t *idx;                  // 8 n stencil cells       (cfdStencil idx)
  float *w;                      // 8 n stencil weights     (cfdStencil w)
  uint8_t *ok;                   // n: stencil found
  double *W;                     // 5 n primitives, NaN where the stencil holds no weight
  double *st;                    // 12 n: P, T, rho, V, M, Tt, Pt, Cp, Taw, Vx, Vy, Vz
  double *vel;                   // 3 n velocity / M, NaN where cfdVelAt has none
  uint8_t *vok;                  // n: velocity found
};

constexpr double kProbeGamma = 1.4;

// Math.max of the reference: NaN in, NaN out
__device__ inline double js_max(double a, double b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }

// cell i below v and the fraction of the way to i + 1 in ascending centres c;
// outside the centres the end interval with fraction 0 or 1 (cfdLocate, src/air_cfd.js:453)
__device__ inline void bracket(const float *__restrict__ c, int n, double v, int &i, double &f) {
#pragma clang fp contract(off)
  if (v <= (double)c[0]) { i = 0; f = 0.0; return; }
  if (v >= (double)c[n - 1]) { i = n - 2; f = 1.0; return; }
  int lo = 0, hi = n - 1;                 // c[lo] <= v < c[hi] (or v NaN)
  while (hi - lo > 1) {
    const int m = (lo + hi) >> 1;
    if ((double)c[m] <= v) lo = m; else hi = m;
  }
  i = lo;
  f = (v - (double)c[lo]) / ((double)c[hi] - (double)c[lo]);
}

// corner weights of the bracket (i, j, k) over cells with a state; corner
// index bit 0 = +x, bit 1 = +y, bit 2 = +z. Returns the FP64 sum of the
// unrounded weights; wt receives them rounded to FP32.
__device__ inline double corner_weights(const ProbeArgs &a, int i, int j, int k, const double fx[2], const double fy[2],
                                        const double fz[2], int q8[8], float wt[8]) {
#pragma clang fp contract(off)
  double ws = 0.0;
  for (int bz = 0; bz < 2; bz++)
    for (int by = 0; by < 2; by++)
      for (int bx = 0; bx < 2; bx++) {
        const int c = bx | by << 1 | bz << 2;
        const int q = ((k + bz) * a.NY + (j + by)) * a.NX + (i + bx);
        const uint8_t tq = a.t[q];
        const double wc = (tq == 0 || tq >= 3) ? fx[bx] * fy[by] * fz[bz] : 0.0;
        q8[c] = q;
        wt[c] = (float)wc;
        ws += wc;
      }
  return ws;
}

// blended primitives of the stencil; false when it holds no weight
__device__ inline bool blend(const ProbeArgs &a, const int q8[8], const float wt[8], double ysgn, double Wp[5]) {
#pragma clang fp contract(off)
  double acc[5] = {0, 0, 0, 0, 0}, ws = 0.0;
  for (int c = 0; c < 8; c++) {
    const double wc = wt[c];
    if (wc == 0.0 || wc != wc) continue;
    for (int v = 0; v < 5; v++) acc[v] += wc * (double)a.Q[v * a.ns + q8[c]];
    ws += wc;
  }
  if (ws == 0.0) return false;
  const double r = js_max(acc[0], 1e-6);
  const double u = acc[1] / r, v = acc[2] / r * ysgn, w = acc[3] / r;
  Wp[0] = r; Wp[1] = u; Wp[2] = v; Wp[3] = w;
  Wp[4] = js_max((kProbeGamma - 1) * (acc[4] - 0.5 * r * (u * u + v * v + w * w)), 1e-7);
  return true;
}

__global__ void __launch_bounds__(256) probe_kernel(ProbeArgs a) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n) return;
  const double x = a.pts[3 * p], y = a.pts[3 * p + 1], z = a.pts[3 * p + 2];
  const double ay = fabs(y), ysgn = y < 0 ? -1.0 : 1.0;
  int i, j, k;
  double tx, ty, tz;
  bracket(a.xc, a.NX, x, i, tx);
  bracket(a.yc, a.NY, ay, j, ty);
  bracket(a.zc, a.NZ, z, k, tz);
  const double fx[2] = {1 - tx, tx}, fy[2] = {1 - ty, ty}, fz[2] = {1 - tz, tz};
  int q8[8];
  float wt[8];
  const double ws = corner_weights(a, i, j, k, fx, fy, fz, q8, wt);

  // ---- velocity (no fallback; inside the box of centres only)
  {
    const bool inside = !(x < (double)a.xc[0] || x > (double)a.xc[a.NX - 1] || ay > (double)a.yc[a.NY - 1] ||
                          z < (double)a.zc[0] || z > (double)a.zc[a.NZ - 1]);
    double Wp[5];
    bool got = false;
    if (inside && !(ws < 1e-6)) {
      float wv[8];
      for (int c = 0; c < 8; c++) wv[c] = (float)((double)wt[c] / ws);
      got = blend(a, q8, wv, ysgn, Wp);
    }
    const double Mi = js_max(a.M, 1e-3);
    if (a.vel)
      for (int c = 0; c < 3; c++) a.vel[3 * p + c] = got ? Wp[1 + c] / Mi : __builtin_nan("");
    if (a.vok) a.vok[p] = got;
  }

  // ---- display stencil, with the nearest fluid centre when the blend holds no weight
  bool found;
  if (ws > 1e-6) {
    for (int c = 0; c < 8; c++) wt[c] = (float)((double)wt[c] / ws);
    found = true;
  } else {
    int best = -1;
    double bd = __builtin_inf();
    for (int dk = -1; dk <= 2; dk++)
      for (int dj = -1; dj <= 2; dj++)
        for (int di = -1; di <= 2; di++) {
          const int ii = i + di, jj = j + dj, kk = k + dk;
          if (ii < 0 || jj < 0 || kk < 0 || ii >= a.NX || jj >= a.NY || kk >= a.NZ) continue;
          const int q = (kk * a.NY + jj) * a.NX + ii;
          if (a.t[q] != 0) continue;
          const double ex = (double)a.xc[ii] - x, ey = (double)a.yc[jj] - ay, ez = (double)a.zc[kk] - z;
          const double d2 = ex * ex + ey * ey + ez * ez;
          if (d2 < bd) { bd = d2; best = q; }
        }
    found = best >= 0;
    for (int c = 0; c < 8; c++) {
      q8[c] = found ? best : 0;
      wt[c] = (found && c == 0) ? 1.0f : 0.0f;
    }
  }
  if (a.idx)
    for (int c = 0; c < 8; c++) a.idx[8 * p + c] = q8[c];
  if (a.w)
    for (int c = 0; c < 8; c++) a.w[8 * p + c] = wt[c];
  if (a.ok) a.ok[p] = found;

  double Wp[5];
  const bool got = blend(a, q8, wt, ysgn, Wp);
  const double nan = __builtin_nan("");
  if (a.W)
    for (int c = 0; c < 5; c++) a.W[5 * p + c] = got ? Wp[c] : nan;
  if (!a.st) return;
  double *o = a.st + 12 * (size_t)p;
  if (!got) {
    for (int c = 0; c < 12; c++) o[c] = nan;
    return;
  }
  // dimensional state: T from p / rho, speed of sound of air (R = 287.05), Mach,
  // isentropic total temperature and pressure, adiabatic wall temperature (recovery 0.89)
  const double T = kProbeGamma * Wp[4] / Wp[0] * a.fsT;
  const double snd = sqrt(1.4 * 287.05 * T);
  const double V = sqrt(Wp[1] * Wp[1] + Wp[2] * Wp[2] + Wp[3] * Wp[3]) * a.fsa;
  const double P = Wp[4] * kProbeGamma * a.fsP, M = V / snd;
  const double m2 = 0.2 * M * M;
  o[0] = P;
  o[1] = T;
  o[2] = Wp[0] * a.fsrho;
  o[3] = V;
  o[4] = M;
  o[5] = T * (1 + m2);
  o[6] = P * pow(1 + m2, 3.5);
  o[7] = a.fsq > 1 ? (P - a.fsP) / a.fsq : 0.0;
  o[8] = T * (1 + 0.89 * 0.2 * M * M);
  o[9] = Wp[1] * a.fsa;
  o[10] = Wp[2] * a.fsa;
  o[11] = Wp[3] * a.fsa;
}

}  // namespace aero3d
