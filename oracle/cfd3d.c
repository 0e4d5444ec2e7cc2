/* ============================================================================
 * oracle/cfd3d.c — CPU restatement of the reference 3D Euler cut-cell solver.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker of the HIP solver and the
 * CPU baseline of the benchmark; no product code path loads it.
 *
 * It restates CfdCPU of the reference (src/air_cfd.js:157-259): JS-double
 * arithmetic, Float32 state storage, Float64 primitives and local time steps,
 * the same loop order and the same operation order inside every expression,
 * so that on identical inputs it reproduces the reference bit for bit (the
 * golden vectors under tests/golden pin this). Post-processing restates
 * cfdForces (src/air_cfd.js:443-452), the readback's Trefftz sum
 * (src/air_cfd.js:560-570), density residual (src/air_cfd.js:553-554) and
 * Mach maximum (src/air_cfd.js:571).
 *
 * The only extension is OpenMP over cells inside a stage (cells are
 * independent within a stage), used to time the baseline on all host cores;
 * it does not change any result.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 * ========================================================================== */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CG 1.4

typedef struct {
  int NX, NY, NZ, N;
  float *dx, *dy, *dz;          /* cell widths, Float32 like the reference grid */
  uint8_t *t;                   /* 0 fluid, 1 solid, 3 inlet sink, 4 bypass jet, 5 core jet */
  float *phi, *ap[4];           /* volume fraction, + face apertures, baffle slope */
  float *Q[5], *B[5], *S1[5];   /* conserved state planes (Float32) */
  double *W[5], *dts;           /* primitives and local dt (Float64) */
  double F[5], fix[3][5];
  double cfl;
  int order, it, threads;
} O3;

/* Math.min / Math.max of two non-NaN doubles including the signed-zero rule. */
static inline double jmin(double a, double b) { if (a < b) return a; if (b < a) return b; return signbit(a) ? a : b; }
static inline double jmax(double a, double b) { if (a > b) return a; if (b > a) return b; return signbit(a) ? b : a; }

/* cfdPrim5: src/air_cfd.js:169 */
static inline void prim5(double a0, double a1, double a2, double a3, double a4, double *o) {
  const double r = jmax(a0, 1e-6), u = a1 / r, v = a2 / r, w = a3 / r;
  o[0] = r; o[1] = u; o[2] = v; o[3] = w;
  o[4] = jmax((CG - 1) * (a4 - 0.5 * r * (u * u + v * v + w * w)), 1e-7);
}

/* cfdHLLC: src/air_cfd.js:179-195. L, R primitives (rho,u,v,w,p); d normal axis. */
static void hllc(const double *L, const double *R, int d, double *out) {
  const double gm = CG - 1, rL = L[0], rR = R[0], pL = L[4], pR = R[4], uL = L[1 + d], uR = R[1 + d];
  const double aL = sqrt(CG * pL / rL), aR = sqrt(CG * pR / rR);
  const double EL = pL / gm + 0.5 * rL * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]);
  const double ER = pR / gm + 0.5 * rR * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3]);
  const double SL = jmin(uL - aL, uR - aR), SR = jmax(uL + aL, uR + aR);
  if (SL >= 0 || SR <= 0) {
    const double *K = SL >= 0 ? L : R;
    const double uK = K[1 + d], E = SL >= 0 ? EL : ER;
    out[0] = K[0] * uK; out[1] = K[0] * K[1] * uK; out[2] = K[0] * K[2] * uK; out[3] = K[0] * K[3] * uK;
    out[1 + d] += K[4];
    out[4] = uK * (E + K[4]);
    return;
  }
  const double SM = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR));
  const double *K; double S, uK, rK, pK, EK;
  if (SM >= 0) { K = L; S = SL; uK = uL; rK = rL; pK = pL; EK = EL; }
  else { K = R; S = SR; uK = uR; rK = rR; pK = pR; EK = ER; }
  const double f = rK * (S - uK) / (S - SM);
  double Fm[4] = { rK * uK, rK * K[1] * uK, rK * K[2] * uK, rK * K[3] * uK };
  Fm[1 + d] += pK;
  const double Um[4] = { rK, rK * K[1], rK * K[2], rK * K[3] };
  double Us[4] = { f, f * K[1], f * K[2], f * K[3] };
  Us[1 + d] = f * SM;
  const double Es = f * (EK / rK + (SM - uK) * (SM + pK / (rK * (S - uK))));
  for (int c = 0; c < 4; c++) out[c] = Fm[c] + S * (Us[c] - Um[c]);
  out[4] = uK * (EK + pK) + S * (Es - EK);
}

static inline double mmod(double a, double b) { return a * b <= 0 ? 0 : (fabs(a) < fabs(b) ? a : b); }

/* CfdCPU.prims: src/air_cfd.js:170-177 */
static void prims(O3 *s, float *const *Q) {
  const int N = s->N;
#pragma omp parallel for schedule(static) num_threads(s->threads)
  for (int q = 0; q < N; q++) {
    double o[5];
    if (s->t[q] >= 3) { const double *f = s->fix[s->t[q] - 3]; prim5(f[0], f[1], f[2], f[3], f[4], o); }
    else prim5(Q[0][q], Q[1][q], Q[2][q], Q[3][q], Q[4][q], o);
    for (int v = 0; v < 5; v++) s->W[v][q] = o[v];
  }
}

/* CfdCPU.cellResidual: src/air_cfd.js:198-236. Residual of fluid cell (i,j,k), divided by phi. */
static void cell_residual(const O3 *s, int i, int j, int k, double *R) {
  const int NX = s->NX, NY = s->NY, NZ = s->NZ, o2 = s->order == 2;
  const int q0 = (k * NY + j) * NX + i;
  double P[25], A[4], sl[15], L[5], Rr[5], fl[5];
  int T[5];
  for (int c = 0; c < 5; c++) R[c] = 0;
  const double pC = s->W[4][q0];
  for (int d = 0; d < 3; d++) {
    const int st = d == 0 ? 1 : d == 1 ? NX : NX * NY, ic = d == 0 ? i : d == 1 ? j : k, nd = d == 0 ? NX : d == 1 ? NY : NZ;
    const double h = d == 0 ? s->dx[i] : d == 1 ? s->dy[j] : s->dz[k];
    const float *apd = s->ap[d];
    for (int m = -2; m <= 2; m++) {
      const int c = ic + m, o = 5 * (m + 2);
      if (c >= nd || (c < 0 && d != 1)) { T[m + 2] = -1; prim5(s->F[0], s->F[1], s->F[2], s->F[3], s->F[4], P + o); continue; }
      if (c < 0) {                                   /* symmetry plane: mirror, v negated */
        const int qm = q0 + (-1 - c - ic) * st;
        T[m + 2] = -2;
        for (int v = 0; v < 5; v++) P[o + v] = s->W[v][qm];
        P[o + 2] = -P[o + 2];
        continue;
      }
      const int q = q0 + m * st;
      T[m + 2] = s->t[q];
      for (int v = 0; v < 5; v++) P[o + v] = s->W[v][q];
    }
    for (int m = -2; m <= 1; m++) { const int c = ic + m; A[m + 2] = (c < 0 || c >= nd) ? 1 : apd[q0 + m * st]; }
    for (int x = -1; x <= 1; x++) {
      const int ok = o2 && T[x + 1] == 0 && T[x + 2] == 0 && T[x + 3] == 0 && A[x + 1] == 1 && A[x + 2] == 1;
      for (int v = 0; v < 5; v++)
        sl[5 * (x + 1) + v] = ok ? mmod(P[5 * (x + 2) + v] - P[5 * (x + 1) + v], P[5 * (x + 3) + v] - P[5 * (x + 2) + v]) : 0;
    }
    for (int sg = 1; sg >= -1; sg -= 2) {
      const double a = sg > 0 ? A[2] : A[1];
      if (a <= 0) continue;
      const int lo = sg > 0 ? 0 : -1;
      for (int v = 0; v < 5; v++) { L[v] = P[5 * (lo + 2) + v] + 0.5 * sl[5 * (lo + 1) + v]; Rr[v] = P[5 * (lo + 3) + v] - 0.5 * sl[5 * (lo + 2) + v]; }
      if (L[0] <= 0 || L[4] <= 0 || Rr[0] <= 0 || Rr[4] <= 0)
        for (int v = 0; v < 5; v++) { L[v] = P[5 * (lo + 2) + v]; Rr[v] = P[5 * (lo + 3) + v]; }
      hllc(L, Rr, d, fl);
      for (int c = 0; c < 5; c++) R[c] -= sg * a * fl[c] / h;
    }
    R[1 + d] += pC * (A[2] - A[1]) / h;
    if (d == 2) {                                    /* baffle transpiration */
      const double sUp = s->ap[3][q0], sDn = k > 0 ? s->ap[3][q0 - NX * NY] : 0;
      if (sUp != 0 || sDn != 0) {
        const double r = s->W[0][q0], u = s->W[1][q0], v = s->W[2][q0], w = s->W[3][q0];
        const double Hh = (pC / (CG - 1) + 0.5 * r * (u * u + v * v + w * w) + pC) / r;
        const double m = (sDn != 0 ? (1 - A[1]) * r * u * sDn : 0) - (sUp != 0 ? (1 - A[2]) * r * u * sUp : 0);
        R[0] += m / h; R[1] += m * u / h; R[2] += m * v / h; R[3] += m * w / h; R[4] += m * Hh / h;
      }
    }
  }
  const double ph = s->phi[q0];
  for (int c = 0; c < 5; c++) R[c] /= ph;
}

/* CfdCPU.stage: src/air_cfd.js:237-251 */
static void stage(O3 *s, float *const *Qin, float *const *Bse, float *const *Out, double wa, double wb, int first) {
  prims(s, Qin);
  const int NX = s->NX, NY = s->NY, NZ = s->NZ;
#pragma omp parallel for schedule(static) num_threads(s->threads)
  for (int kj = 0; kj < NZ * NY; kj++) {
    const int k = kj / NY, j = kj % NY;
    double R[5];
    for (int i = 0; i < NX; i++) {
      const int q = (k * NY + j) * NX + i, tq = s->t[q];
      if (tq != 0) {
        if (tq >= 3) { const double *f = s->fix[tq - 3]; for (int c = 0; c < 5; c++) Out[c][q] = (float)f[c]; }
        else for (int c = 0; c < 5; c++) Out[c][q] = Bse[c][q];
        continue;
      }
      if (first) {
        const double a = sqrt(CG * s->W[4][q] / s->W[0][q]);
        s->dts[q] = s->cfl * jmax(s->phi[q], 0.25) / ((fabs(s->W[1][q]) + a) / s->dx[i] + (fabs(s->W[2][q]) + a) / s->dy[j] + (fabs(s->W[3][q]) + a) / s->dz[k]);
      }
      cell_residual(s, i, j, k, R);
      const double dt = s->dts[q];
      for (int c = 0; c < 5; c++) Out[c][q] = (float)(wa * Bse[c][q] + wb * (Qin[c][q] + dt * R[c]));
      const double r = jmax(Out[0][q], 1e-4);
      Out[0][q] = (float)r;
      const double o1 = Out[1][q], o2 = Out[2][q], o3 = Out[3][q];
      const double ke = 0.5 * (o1 * o1 + o2 * o2 + o3 * o3) / r;
      if ((CG - 1) * (Out[4][q] - ke) < 1e-5) Out[4][q] = (float)(1e-5 / (CG - 1) + ke);
    }
  }
}

/* ---------------- C interface (ctypes) ---------------- */
static float *dupf(const float *p, int n) { float *d = (float *)malloc(sizeof(float) * (size_t)n); memcpy(d, p, sizeof(float) * (size_t)n); return d; }

O3 *o3_create(int NX, int NY, int NZ, const float *dx, const float *dy, const float *dz, const uint8_t *t,
              const float *phi, const float *ap0, const float *ap1, const float *ap2, const float *ap3) {
  /* the reference's y = 0 mirror reads rows 0 and 1 (src/air_cfd.js:205-210) */
  if (NX < 1 || NY < 2 || NZ < 1 || (long long)NX * NY * NZ > (1LL << 30)) return NULL;
  O3 *s = (O3 *)calloc(1, sizeof(O3));
  s->NX = NX; s->NY = NY; s->NZ = NZ; s->N = NX * NY * NZ;
  const int N = s->N;
  s->dx = dupf(dx, NX); s->dy = dupf(dy, NY); s->dz = dupf(dz, NZ);
  s->t = (uint8_t *)malloc((size_t)N); memcpy(s->t, t, (size_t)N);
  s->phi = dupf(phi, N); s->ap[0] = dupf(ap0, N); s->ap[1] = dupf(ap1, N); s->ap[2] = dupf(ap2, N); s->ap[3] = dupf(ap3, N);
  for (int v = 0; v < 5; v++) {
    s->Q[v] = (float *)calloc((size_t)N, 4); s->B[v] = (float *)calloc((size_t)N, 4); s->S1[v] = (float *)calloc((size_t)N, 4);
    s->W[v] = (double *)calloc((size_t)N, 8);
  }
  s->dts = (double *)calloc((size_t)N, 8);
  s->cfl = 0.7; s->order = 2; s->threads = 1;
  const double F[5] = { 1, 0, 0, 0, 1 / (CG * (CG - 1)) };
  memcpy(s->F, F, sizeof F);
  for (int g = 0; g < 3; g++) memcpy(s->fix[g], F, sizeof F);
  return s;
}
void o3_destroy(O3 *s) {
  if (!s) return;
  free(s->dx); free(s->dy); free(s->dz); free(s->t); free(s->phi);
  for (int d = 0; d < 4; d++) free(s->ap[d]);
  for (int v = 0; v < 5; v++) { free(s->Q[v]); free(s->B[v]); free(s->S1[v]); free(s->W[v]); }
  free(s->dts); free(s);
}
void o3_set_free(O3 *s, const double *F) { memcpy(s->F, F, 5 * sizeof(double)); }
void o3_set_fix(O3 *s, const double *fix) { memcpy(s->fix, fix, 15 * sizeof(double)); }
void o3_set_opt(O3 *s, double cfl, int order) { s->cfl = cfl; s->order = order; }
void o3_set_threads(O3 *s, int n) { s->threads = n > 0 ? n : 1; }
int o3_iter(const O3 *s) { return s->it; }
/* CfdCPU.reset: src/air_cfd.js:168 */
void o3_reset(O3 *s) {
  for (int q = 0; q < s->N; q++) for (int c = 0; c < 5; c++) s->Q[c][q] = (float)s->F[c];
  s->it = 0;
}
/* CfdCPU.step: src/air_cfd.js:252-258 (SSP-RK2, Shu-Osher weights) */
void o3_step(O3 *s, int n) {
  for (int k = 0; k < n; k++) {
    stage(s, s->Q, s->Q, s->S1, 0, 1, 1);
    stage(s, s->S1, s->Q, s->B, 0.5, 0.5, 0);
    for (int v = 0; v < 5; v++) { float *tmp = s->Q[v]; s->Q[v] = s->B[v]; s->B[v] = tmp; }
    s->it++;
  }
}
void o3_read(const O3 *s, float *r, float *ru, float *rv, float *rw, float *E) {
  const size_t b = sizeof(float) * (size_t)s->N;
  memcpy(r, s->Q[0], b); memcpy(ru, s->Q[1], b); memcpy(rv, s->Q[2], b); memcpy(rw, s->Q[3], b); memcpy(E, s->Q[4], b);
}
void o3_write(O3 *s, const float *r, const float *ru, const float *rv, const float *rw, const float *E) {
  const size_t b = sizeof(float) * (size_t)s->N;
  memcpy(s->Q[0], r, b); memcpy(s->Q[1], ru, b); memcpy(s->Q[2], rv, b); memcpy(s->Q[3], rw, b); memcpy(s->Q[4], E, b);
}
/* Local dt of the last first stage (Float64). */
void o3_read_dt(const O3 *s, double *dt) { memcpy(dt, s->dts, sizeof(double) * (size_t)s->N); }

/* cfdForces: src/air_cfd.js:443-452. out = {CL, CD, Cm}. */
void o3_forces(const float *r_, const float *ru_, const float *rv_, const float *rw_, const float *E_,
               const int *q_, const double *cx, const double *cz, const double *x, const double *z, int n,
               double alphaDeg, double M, double S, double xcg, double zcg, double cref, double *out) {
  double Fx = 0, Fz = 0, My = 0;
  const double pinf = 1 / CG;
  for (int s = 0; s < n; s++) {
    const int q = q_[s];
    const double r = r_[q], u = ru_[q] / r, v = rv_[q] / r, w = rw_[q] / r;
    const double dp = (CG - 1) * (E_[q] - 0.5 * r * (u * u + v * v + w * w)) - pinf;
    const double fx = -dp * cx[s], fz = -dp * cz[s];
    Fx += fx; Fz += fz; My += (z[s] - zcg) * fx - (x[s] - xcg) * fz;
  }
  const double a = alphaDeg * (M_PI / 180), qd = 0.5 * M * M, k = 2 / (qd * S);
  out[0] = k * (Fz * cos(a) - Fx * sin(a));
  out[1] = k * (Fx * cos(a) + Fz * sin(a));
  out[2] = k * My / cref;
}

/* Trefftz sum of the readback: src/air_cfd.js:561-569. xc: Float32 x centres. */
double o3_trefftz(int NX, int NY, int NZ, const float *xc, const float *dy, const float *dz, const uint8_t *t,
                  const float *r_, const float *ru_, const float *rv_, const float *rw_, double xT, double alphaDeg, double M, double S) {
  int i0; double tx;
  if (xT <= xc[0]) { i0 = 0; tx = 0; }
  else if (xT >= xc[NX - 1]) { i0 = NX - 2; tx = 1; }
  else {
    int lo = 0, hi = NX - 1;
    while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (xc[m] <= xT) lo = m; else hi = m; }
    i0 = lo; tx = (xT - xc[lo]) / ((double)xc[hi] - xc[lo]);
  }
  const double a = alphaDeg * (M_PI / 180), ca = cos(a), sa = sin(a), Mn = jmax(M, 0.05);
  double sum = 0;
  for (int k = 0; k < NZ; k++) for (int j = 0; j < NY; j++) {
    double u = 0, v = 0, w = 0, ws = 0;
    for (int di = 0; di < 2; di++) {
      const int q = (k * NY + j) * NX + i0 + di;
      if (t[q] != 0) continue;
      const double ww = di ? tx : 1 - tx, r = r_[q];
      u += ww * ru_[q] / r; v += ww * rv_[q] / r; w += ww * rw_[q] / r; ws += ww;
    }
    if (ws < 1e-6) continue;
    u /= ws; v /= ws; w /= ws;
    const double vn = w * ca - u * sa;
    sum += (v * v + vn * vn) * dy[j] * dz[k];
  }
  return 2 * sum / (Mn * Mn * S);
}

/* Density residual between reads: src/air_cfd.js:553-554. */
double o3_residual(int N, const uint8_t *t, const float *r, const float *prev, int nIt) {
  double sum = 0; int n = 0;
  for (int q = 0; q < N; q += 3) { if (t[q] != 0) continue; const double d = (double)r[q] - prev[q]; sum += d * d; n++; }
  return sqrt(sum / (n > 1 ? n : 1)) / nIt;
}

/* Mach maximum of the readback: src/air_cfd.js:571. Every second cell, fluid
 * cells only: m2 = (u^2 + v^2 + w^2) / (CG p / r) in doubles from the Float32
 * state; the square root of the largest m2 (0 with no fluid cell; a NaN m2
 * never compares larger). */
double o3_mach_max(int N, const uint8_t *t, const float *r_, const float *ru_, const float *rv_, const float *rw_, const float *E_) {
  double mMax = 0;
  for (int q = 0; q < N; q += 2) {
    if (t[q] != 0) continue;
    const double r = r_[q], u = ru_[q] / r, v = rv_[q] / r, w = rw_[q] / r;
    const double p = (CG - 1) * (E_[q] - 0.5 * r * (u * u + v * v + w * w));
    const double m2 = (u * u + v * v + w * w) / (CG * p / r);
    if (m2 > mMax) mMax = m2;
  }
  return sqrt(mMax);
}
