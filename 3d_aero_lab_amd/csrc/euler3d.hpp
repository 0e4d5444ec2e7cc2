// ============================================================================
// euler3d.hpp — device arithmetic of one SSP-RK2 stage of the 3D Euler
// cut-cell scheme, FP32, shared by the plain and batched kernels.
//
// Scheme (reference: CfdCPU, src/air_cfd.js:169-251; CFD3_FS, src/air_cfd.js:266-366):
//   primitives  rho = max(rho,1e-6), p = max((g-1)(E - rho|u|^2/2), 1e-7)
//   per axis d  gather cells ic-2..ic+2 (freestream ghost outside the box,
//               mirror with v negated below the symmetry plane y = 0),
//               minmod slopes of primitives where three cells are fluid and
//               both faces fully open, MUSCL face states with a first-order
//               fallback on non-positive rho or p, HLLC flux with Davis
//               speeds, weighted by the face aperture;
//               wall pressure source p_C (A+ - A-) / h; on z, the baffle
//               transpiration mass m (1, u, v, w, H) / h
//   update      R / phi, out = wa * base + wb * (in + dt R), floors
//               rho >= 1e-4, p >= 1e-5; dt = cfl max(phi, 0.25) /
//               sum_d (|u_d| + a) / h_d from the base state.
// ============================================================================
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aero3d {

constexpr float kGamma = 1.4f;
constexpr float kGm1 = 0.4f;   // (float)(1.4 - 1)

// Per-solve constants of a stage.
struct StageConsts {
  float Fw[5];        // freestream primitives (rho, u, v, w, p)
  float fixw[3][5];   // primitives of fixed states, flags 3..5
  float fixq[3][5];   // conserved fixed states, flags 3..5
  float cfl;
  int order;
};

// Device view of one solve: grid, masks, and the three state buffers of a stage.
struct CaseView {
  int NX, NY, NZ, N;
  const float *__restrict__ dx, *__restrict__ dy, *__restrict__ dz;
  const uint8_t *__restrict__ t;
  const float *__restrict__ phi;
  const float *__restrict__ ap;   // 4 planes of N: ap0, ap1, ap2, ap3
};

__device__ __forceinline__ void prim5(float a0, float a1, float a2, float a3, float a4, float *o) {
  const float r = fmaxf(a0, 1e-6f), u = a1 / r, v = a2 / r, w = a3 / r;
  o[0] = r; o[1] = u; o[2] = v; o[3] = w;
  o[4] = fmaxf(kGm1 * (a4 - 0.5f * r * (u * u + v * v + w * w)), 1e-7f);
}

// HLLC flux along axis d of primitive states L, R (src/air_cfd.js:179-195).
template <int d>
__device__ __forceinline__ void hllc(const float *L, const float *R, float *out) {
  const float rL = L[0], rR = R[0], pL = L[4], pR = R[4], uL = L[1 + d], uR = R[1 + d];
  const float aL = sqrtf(kGamma * pL / rL), aR = sqrtf(kGamma * pR / rR);
  const float EL = pL / kGm1 + 0.5f * rL * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]);
  const float ER = pR / kGm1 + 0.5f * rR * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3]);
  const float SL = fminf(uL - aL, uR - aR), SR = fmaxf(uL + aL, uR + aR);
  if (SL >= 0.f || SR <= 0.f) {
    const bool left = SL >= 0.f;
    const float *K = left ? L : R;
    const float uK = K[1 + d], E = left ? EL : ER;
    out[0] = K[0] * uK; out[1] = K[0] * K[1] * uK; out[2] = K[0] * K[2] * uK; out[3] = K[0] * K[3] * uK;
    out[1 + d] += K[4];
    out[4] = uK * (E + K[4]);
    return;
  }
  const float SM = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR));
  const bool left = SM >= 0.f;
  const float *K = left ? L : R;
  const float S = left ? SL : SR, uK = left ? uL : uR, rK = left ? rL : rR, pK = left ? pL : pR, EK = left ? EL : ER;
  const float f = rK * (S - uK) / (S - SM);
  float Fm[4] = {rK * uK, rK * K[1] * uK, rK * K[2] * uK, rK * K[3] * uK};
  Fm[1 + d] += pK;
  const float Um[4] = {rK, rK * K[1], rK * K[2], rK * K[3]};
  float Us[4] = {f, f * K[1], f * K[2], f * K[3]};
  Us[1 + d] = f * SM;
  const float Es = f * (EK / rK + (SM - uK) * (SM + pK / (rK * (S - uK))));
#pragma unroll
  for (int c = 0; c < 4; c++) out[c] = Fm[c] + S * (Us[c] - Um[c]);
  out[4] = uK * (EK + pK) + S * (Es - EK);
}

__device__ __forceinline__ float mmod(float a, float b) { return a * b <= 0.f ? 0.f : (fabsf(a) < fabsf(b) ? a : b); }

// Accumulate the axis-d contribution of fluid cell (i,j,k) into R
// (src/air_cfd.js:203-234): stencil gather, slopes, two aperture-weighted
// HLLC fluxes, wall pressure source and, for d = 2, baffle transpiration.
template <int d>
__device__ __forceinline__ void axis_residual(const CaseView &cv, const StageConsts &kc, const float *__restrict__ qin,
                                              size_t ns, int i, int j, int k, int q0, const float *C, float pC,
                                              float *R) {
  const int NX = cv.NX, NY = cv.NY;
  const int st = d == 0 ? 1 : d == 1 ? NX : NX * NY;
  const int ic = d == 0 ? i : d == 1 ? j : k;
  const int nd = d == 0 ? cv.NX : d == 1 ? cv.NY : cv.NZ;
  const float h = d == 0 ? cv.dx[i] : d == 1 ? cv.dy[j] : cv.dz[k];
  const float *__restrict__ apd = cv.ap + (size_t)d * cv.N;
  float P[5][5];
  int T[5];
#pragma unroll
  for (int m = -2; m <= 2; m++) {
    const int c = ic + m;
    float *o = P[m + 2];
    if (m == 0) {
      T[2] = 0;
#pragma unroll
      for (int v = 0; v < 5; v++) o[v] = C[v];
      o[4] = pC;
      continue;
    }
    if (c >= nd || (c < 0 && d != 1)) {
      T[m + 2] = -1;
#pragma unroll
      for (int v = 0; v < 5; v++) o[v] = kc.Fw[v];
      continue;
    }
    int q;
    bool mirror = false;
    if (c < 0) { q = q0 + (-1 - c - ic) * st; mirror = true; T[m + 2] = -2; }
    else { q = q0 + m * st; }
    const int tq = cv.t[q];
    if (!mirror) T[m + 2] = tq;
    if (tq >= 3) {
#pragma unroll
      for (int v = 0; v < 5; v++) o[v] = kc.fixw[tq - 3][v];
    } else {
      prim5(qin[q], qin[ns + q], qin[2 * ns + q], qin[3 * ns + q], qin[4 * ns + q], o);
    }
    if (mirror) o[2] = -o[2];
  }
  float A[4];
#pragma unroll
  for (int m = -2; m <= 1; m++) {
    const int c = ic + m;
    A[m + 2] = (c < 0 || c >= nd) ? 1.f : apd[q0 + m * st];
  }
  float s[3][5];
  const bool o2 = kc.order == 2;
#pragma unroll
  for (int x = 0; x < 3; x++) {
    const bool ok = o2 && T[x] == 0 && T[x + 1] == 0 && T[x + 2] == 0 && A[x] == 1.f && A[x + 1] == 1.f;
#pragma unroll
    for (int v = 0; v < 5; v++) s[x][v] = ok ? mmod(P[x + 1][v] - P[x][v], P[x + 2][v] - P[x + 1][v]) : 0.f;
  }
#pragma unroll
  for (int side = 0; side < 2; side++) {           // side 0: + face (sg = +1), side 1: - face (sg = -1)
    const float a = side == 0 ? A[2] : A[1];
    if (a <= 0.f) continue;
    const int lo = side == 0 ? 2 : 1;               // index of the face's left cell in P
    float L[5], Rr[5], fl[5];
#pragma unroll
    for (int v = 0; v < 5; v++) { L[v] = P[lo][v] + 0.5f * s[lo - 1][v]; Rr[v] = P[lo + 1][v] - 0.5f * s[lo][v]; }
    if (L[0] <= 0.f || L[4] <= 0.f || Rr[0] <= 0.f || Rr[4] <= 0.f) {
#pragma unroll
      for (int v = 0; v < 5; v++) { L[v] = P[lo][v]; Rr[v] = P[lo + 1][v]; }
    }
    hllc<d>(L, Rr, fl);
    const float sa = side == 0 ? a : -a;
#pragma unroll
    for (int c = 0; c < 5; c++) R[c] -= sa * fl[c] / h;
  }
  R[1 + d] += pC * (A[2] - A[1]) / h;
  if (d == 2) {
    const float *__restrict__ ap3 = cv.ap + 3 * (size_t)cv.N;
    const float sUp = ap3[q0], sDn = k > 0 ? ap3[q0 - NX * NY] : 0.f;
    if (sUp != 0.f || sDn != 0.f) {
      const float r = C[0], u = C[1], v = C[2], w = C[3];
      const float Hh = (pC / kGm1 + 0.5f * r * (u * u + v * v + w * w) + pC) / r;
      const float m = (sDn != 0.f ? (1.f - A[1]) * r * u * sDn : 0.f) - (sUp != 0.f ? (1.f - A[2]) * r * u * sUp : 0.f);
      R[0] += m / h; R[1] += m * u / h; R[2] += m * v / h; R[3] += m * w / h; R[4] += m * Hh / h;
    }
  }
}

// One stage for cell q (src/air_cfd.js:237-251 with dt recomputed from the
// base state as CFD3_FS does, src/air_cfd.js:357-359). Writes out[], and
// dtout[q] when dtout is non-null.
__device__ __forceinline__ void stage_cell(const CaseView &cv, const StageConsts &kc, const float *__restrict__ qin,
                                           const float *__restrict__ bse, float *__restrict__ out,
                                           float *__restrict__ dtout, size_t ns, float wa, float wb, int q) {
  const int NX = cv.NX, NY = cv.NY;
  const int i = q % NX, jk = q / NX, j = jk % NY, k = jk / NY;
  const int tq = cv.t[q];
  if (tq != 0) {
    if (tq >= 3) {
#pragma unroll
      for (int c = 0; c < 5; c++) out[c * ns + q] = kc.fixq[tq - 3][c];
    } else {
#pragma unroll
      for (int c = 0; c < 5; c++) out[c * ns + q] = bse[c * ns + q];
    }
    if (dtout) dtout[q] = 0.f;
    return;
  }
  float Qc[5], C[5], Bq[5], Bw[5];
#pragma unroll
  for (int c = 0; c < 5; c++) Qc[c] = qin[c * ns + q];
#pragma unroll
  for (int c = 0; c < 5; c++) Bq[c] = bse[c * ns + q];
  prim5(Qc[0], Qc[1], Qc[2], Qc[3], Qc[4], C);
  const float pC = C[4];
  float R[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  axis_residual<0>(cv, kc, qin, ns, i, j, k, q, C, pC, R);
  axis_residual<1>(cv, kc, qin, ns, i, j, k, q, C, pC, R);
  axis_residual<2>(cv, kc, qin, ns, i, j, k, q, C, pC, R);
  const float ph = cv.phi[q];
#pragma unroll
  for (int c = 0; c < 5; c++) R[c] /= ph;
  prim5(Bq[0], Bq[1], Bq[2], Bq[3], Bq[4], Bw);
  const float a = sqrtf(kGamma * Bw[4] / Bw[0]);
  const float dt = kc.cfl * fmaxf(ph, 0.25f) /
                   ((fabsf(Bw[1]) + a) / cv.dx[i] + (fabsf(Bw[2]) + a) / cv.dy[j] + (fabsf(Bw[3]) + a) / cv.dz[k]);
  float n[5];
#pragma unroll
  for (int c = 0; c < 5; c++) n[c] = wa * Bq[c] + wb * (Qc[c] + dt * R[c]);
  n[0] = fmaxf(n[0], 1e-4f);
  const float ke = 0.5f * (n[1] * n[1] + n[2] * n[2] + n[3] * n[3]) / n[0];
  if (kGm1 * (n[4] - ke) < 1e-5f) n[4] = 1e-5f / kGm1 + ke;
#pragma unroll
  for (int c = 0; c < 5; c++) out[c * ns + q] = n[c];
  if (dtout) dtout[q] = dt;
}

}  // namespace aero3d
