/* ============================================================================
 * oracle/cfd_axi.c — CPU restatement of the reference axisymmetric engine
 * solver (CPUSolver, src/cfd_cpu.js:7-210).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP axisymmetric solver and the
 * CPU baseline of its benchmark; no product code path loads it.
 *
 * Float64 state and arithmetic as the reference, the same face loops (x faces
 * by row, r faces by face row), the same operation order, Math.min/max and
 * clamp semantics. Zones arrive as the reference's typed arrays (Float32 /
 * Uint8) and are widened to double where the reference reads them.
 * ========================================================================== */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define RGAS 287.05

typedef struct {
  int W, H, N;
  double *xc, *rc, *dx, *dr, *rf;
  uint8_t *solid, *comp;
  float *phi, *ae, *an, *wgt, *lmix, *tgt, *rate, *yrate, *mfd;
  double comps[32];
  double fs_r, fs_u, fs_p;
  double cfl, hf, dtRatio;
  int order, visc, global;
  double *Q[5], *Q1[5], *R[5], *dt, *nu, *src, *mfx;
  double *Pr, *Pu, *Pv, *Pp, *PY, *Pg;
  double *Fx[5], *Fr[5];
  int iter;
} AX;

static inline double jmin(double a, double b) { if (a < b) return a; if (b < a) return b; return signbit(a) ? a : b; }
static inline double jmax(double a, double b) { if (a > b) return a; if (b > a) return b; return signbit(a) ? b : a; }
static inline double clampd(double v, double a, double b) { return v < a ? a : v > b ? b : v; }   /* src/field.js:4 */
static inline double minmod(double a, double b) { return a * b <= 0 ? 0 : (fabs(a) < fabs(b) ? a : b); }

/* hllc: src/cfd_cpu.js:7-31 (normal velocity u, tangential v, species Y, per-side gamma) */
static void hllc(double rL, double uL, double vL, double pL, double YL, double gL, double rR, double uR, double vR, double pR,
                 double YR, double gR, double *F) {
  const double cL = sqrt(gL * pL / rL), cR = sqrt(gR * pR / rR);
  const double SL = jmin(uL - cL, uR - cR), SR = jmax(uL + cL, uR + cR);
  const double EL = pL / (gL - 1) + 0.5 * rL * (uL * uL + vL * vL), ER = pR / (gR - 1) + 0.5 * rR * (uR * uR + vR * vR);
  if (SL >= 0) { F[0] = rL * uL; F[1] = rL * uL * uL + pL; F[2] = rL * uL * vL; F[3] = uL * (EL + pL); F[4] = rL * uL * YL; return; }
  if (SR <= 0) { F[0] = rR * uR; F[1] = rR * uR * uR + pR; F[2] = rR * uR * vR; F[3] = uR * (ER + pR); F[4] = rR * uR * YR; return; }
  const double SM = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR));
  if (SM >= 0) {
    const double f = rL * (SL - uL) / (SL - SM);
    const double Es = f * (EL / rL + (SM - uL) * (SM + pL / (rL * (SL - uL))));
    F[0] = rL * uL + SL * (f - rL);
    F[1] = rL * uL * uL + pL + SL * (f * SM - rL * uL);
    F[2] = rL * uL * vL + SL * (f * vL - rL * vL);
    F[3] = uL * (EL + pL) + SL * (Es - EL);
    F[4] = rL * uL * YL + SL * (f * YL - rL * YL);
  } else {
    const double f = rR * (SR - uR) / (SR - SM);
    const double Es = f * (ER / rR + (SM - uR) * (SM + pR / (rR * (SR - uR))));
    F[0] = rR * uR + SR * (f - rR);
    F[1] = rR * uR * uR + pR + SR * (f * SM - rR * uR);
    F[2] = rR * uR * vR + SR * (f * vR - rR * vR);
    F[3] = uR * (ER + pR) + SR * (Es - ER);
    F[4] = rR * uR * YR + SR * (f * YR - rR * YR);
  }
}

/* farfield: Riemann-invariant ghost state, outward normal (nx, nr): src/cfd_cpu.js:35-48 */
static void farfield(const AX *s, double r, double u, double v, double p, double Y, double g, double nx, double nr, double *G) {
  const double un = u * nx + v * nr, c = sqrt(g * p / r);
  const double g0 = 1.4, c0 = sqrt(g0 * s->fs_p / s->fs_r), un0 = s->fs_u * nx;
  if (un >= c) { G[0] = r; G[1] = u; G[2] = v; G[3] = p; G[4] = Y; return; }
  if (un0 <= -c0) { G[0] = s->fs_r; G[1] = s->fs_u; G[2] = 0; G[3] = s->fs_p; G[4] = 0; return; }
  const double Rp = un + 2 * c / (g - 1), Rm = un0 - 2 * c0 / (g0 - 1);
  const double unb = 0.5 * (Rp + Rm);
  double gb, sv, ut, vt, Yb;
  if (unb > 0) { gb = g; sv = p / pow(r, g); ut = u - un * nx; vt = v - un * nr; Yb = Y; }
  else { gb = g0; sv = s->fs_p / pow(s->fs_r, g0); ut = s->fs_u - un0 * nx; vt = -un0 * nr; Yb = 0; }
  const double cb = jmax(1, 0.25 * (gb - 1) * (Rp - Rm));
  const double rb = pow(cb * cb / (gb * sv), 1 / (gb - 1));
  G[0] = rb; G[1] = ut + unb * nx; G[2] = vt + unb * nr; G[3] = rb * cb * cb / gb; G[4] = Yb;
}

/* CPUSolver.prim: src/cfd_cpu.js:67-74 */
static void prim(AX *s, double *const *Q) {
  for (int k = 0; k < s->N; k++) {
    const double r = jmax(Q[0][k], 1e-4), u = Q[1][k] / r, v = Q[2][k] / r, Y = clampd(Q[4][k] / r, 0, 1), g = 1.4 - 0.07 * Y;
    s->Pr[k] = r; s->Pu[k] = u; s->Pv[k] = v; s->PY[k] = Y; s->Pg[k] = g;
    s->Pp[k] = jmax(1, (g - 1) * (Q[3][k] - 0.5 * r * (u * u + v * v)));
  }
}

static inline void sel(const AX *s, int k, int dir, double *o) {
  o[0] = s->Pr[k]; o[1] = dir ? s->Pv[k] : s->Pu[k]; o[2] = dir ? s->Pu[k] : s->Pv[k]; o[3] = s->Pp[k]; o[4] = s->PY[k];
}
/* recon: src/cfd_cpu.js:86-94 */
static void recon(const AX *s, int a, int am, int b, int bp, int dir, double *L, double *Rr) {
  double A5[5], B5[5], T5[5];
  sel(s, a, dir, A5); sel(s, b, dir, B5);
  for (int q = 0; q < 5; q++) { L[q] = A5[q]; Rr[q] = B5[q]; }
  if (s->order != 2) return;
  if (am >= 0) { sel(s, am, dir, T5); for (int q = 0; q < 5; q++) L[q] = A5[q] + 0.5 * minmod(A5[q] - T5[q], B5[q] - A5[q]); }
  if (bp >= 0) { sel(s, bp, dir, T5); for (int q = 0; q < 5; q++) Rr[q] = B5[q] - 0.5 * minmod(B5[q] - A5[q], T5[q] - B5[q]); }
  if (L[0] <= 0 || L[3] <= 0) for (int q = 0; q < 5; q++) L[q] = A5[q];
  if (Rr[0] <= 0 || Rr[3] <= 0) for (int q = 0; q < 5; q++) Rr[q] = B5[q];
}

#define FULL(k) (s->phi[k] > 0.999)

/* CPUSolver.residual: src/cfd_cpu.js:75-186 */
static void residual(AX *s, double *const *Q, int computeDt) {
  const int W = s->W, H = s->H;
  prim(s, Q);
  double F[5], G[5], G2[5], L[5], Rr[5];
  const uint8_t *S = s->solid;
  memset(s->nu, 0, sizeof(double) * (size_t)s->N);
  /* x faces */
  for (int j = 0; j < H; j++) for (int f = 0; f <= W; f++) {
    const int fi = j * (W + 1) + f, a = f - 1, b = f, ka = j * W + a, kb = j * W + b;
    double al;
    if (a < 0) al = S[kb] ? 0 : 1; else if (b >= W) al = S[ka] ? 0 : 1; else al = s->ae[ka];
    if (al <= 0) { for (int q = 0; q < 5; q++) s->Fx[q][fi] = 0; continue; }
    if (a < 0) {
      farfield(s, s->Pr[kb], s->Pu[kb], s->Pv[kb], s->Pp[kb], s->PY[kb], s->Pg[kb], -1, 0, G);
      hllc(G[0], G[1], G[2], G[3], G[4], 1.4 - 0.07 * G[4], s->Pr[kb], s->Pu[kb], s->Pv[kb], s->Pp[kb], s->PY[kb], s->Pg[kb], F);
    } else if (b >= W) {
      farfield(s, s->Pr[ka], s->Pu[ka], s->Pv[ka], s->Pp[ka], s->PY[ka], s->Pg[ka], 1, 0, G);
      hllc(s->Pr[ka], s->Pu[ka], s->Pv[ka], s->Pp[ka], s->PY[ka], s->Pg[ka], G[0], G[1], G[2], G[3], G[4], 1.4 - 0.07 * G[4], F);
    } else {
      const int fa = FULL(ka) && FULL(kb);
      const int am = fa && a - 1 >= 0 && FULL(ka - 1) ? ka - 1 : -1, bp = fa && b + 1 < W && FULL(kb + 1) ? kb + 1 : -1;
      recon(s, ka, am, kb, bp, 0, L, Rr);
      hllc(L[0], L[1], L[2], L[3], L[4], 1.4 - 0.07 * L[4], Rr[0], Rr[1], Rr[2], Rr[3], Rr[4], 1.4 - 0.07 * Rr[4], F);
    }
    for (int q = 0; q < 5; q++) s->Fx[q][fi] = al * F[q];
  }
  /* r faces (rotated: normal v, tangential u) */
  for (int f = 0; f <= H; f++) for (int i = 0; i < W; i++) {
    const int fi = f * W + i, a = f - 1, b = f, ka = a * W + i, kb = b * W + i;
    for (int q = 0; q < 5; q++) G2[q] = 0;
    const double al = a < 0 ? 0 : b >= H ? (S[ka] ? 0 : 1) : (double)s->an[ka];
    if (al <= 0) { for (int q = 0; q < 5; q++) s->Fr[q][fi] = 0; continue; }
    if (b >= H) {
      farfield(s, s->Pr[ka], s->Pu[ka], s->Pv[ka], s->Pp[ka], s->PY[ka], s->Pg[ka], 0, 1, G);
      hllc(s->Pr[ka], s->Pv[ka], s->Pu[ka], s->Pp[ka], s->PY[ka], s->Pg[ka], G[0], G[2], G[1], G[3], G[4], 1.4 - 0.07 * G[4], F);
    } else {
      const int fa = FULL(ka) && FULL(kb);
      const int am = fa && a - 1 >= 0 && FULL(ka - W) ? ka - W : -1, bp = fa && b + 1 < H && FULL(kb + W) ? kb + W : -1;
      recon(s, ka, am, kb, bp, 1, L, Rr);
      hllc(L[0], L[1], L[2], L[3], L[4], 1.4 - 0.07 * L[4], Rr[0], Rr[1], Rr[2], Rr[3], Rr[4], 1.4 - 0.07 * Rr[4], F);
      if (s->visc) {
        const double drc = s->rc[b] - s->rc[a];
        const double dudr = (s->Pu[kb] - s->Pu[ka]) / drc;
        const double l = 0.5 * ((double)s->lmix[ka] + (double)s->lmix[kb]);
        const double nut = l * l * fabs(dudr), rho = 0.5 * (s->Pr[ka] + s->Pr[kb]);
        const double mu = rho * nut;
        const double Ta = s->Pp[ka] / (s->Pr[ka] * RGAS), Tb = s->Pp[kb] / (s->Pr[kb] * RGAS);
        const double ga = 0.5 * (s->Pg[ka] + s->Pg[kb]), cp = ga * RGAS / (ga - 1);
        const double tau = mu * dudr;
        G2[1] = tau;
        G2[3] = 0.5 * (s->Pu[ka] + s->Pu[kb]) * tau + mu * cp / 0.9 * (Tb - Ta) / drc;
        G2[4] = mu / 0.9 * (s->PY[kb] - s->PY[ka]) / drc;
        if (nut > s->nu[ka]) s->nu[ka] = nut;
        if (nut > s->nu[kb]) s->nu[kb] = nut;
      }
    }
    s->Fr[0][fi] = al * F[0]; s->Fr[1][fi] = al * (F[2] - G2[1]); s->Fr[2][fi] = al * F[1];
    s->Fr[3][fi] = al * (F[3] - G2[3]); s->Fr[4][fi] = al * (F[4] - G2[4]);
  }
  /* assembly and sources */
  const double *cmp = s->comps;
  for (int j = 0; j < H; j++) for (int i = 0; i < W; i++) {
    const int k = j * W + i;
    if (S[k]) { for (int q = 0; q < 5; q++) s->R[q][k] = 0; continue; }
    const int fx0 = j * (W + 1) + i, fx1 = fx0 + 1, fr0 = j * W + i, fr1 = fr0 + W;
    const double ax0 = s->rf[j], ax1 = s->rf[j + 1], vol = s->rc[j] * s->dr[j], ph = s->phi[k];
    for (int q = 0; q < 5; q++)
      s->R[q][k] = (-(s->Fx[q][fx1] - s->Fx[q][fx0]) / s->dx[i] - (s->Fr[q][fr1] * ax1 - s->Fr[q][fr0] * ax0) / vol) / ph;
    const double aw = i > 0 ? (double)s->ae[k - 1] : 1, as = j > 0 ? (double)s->an[k - W] : 0;
    s->R[1][k] += s->Pp[k] * ((double)s->ae[k] - aw) / (s->dx[i] * ph);
    s->R[2][k] += s->Pp[k] * ((double)s->an[k] * ax1 - as * ax0) / (vol * ph);
    const double mfx = 0.5 * (s->Fx[0][fx0] + s->Fx[0][fx1]);
    s->mfx[k] = mfx; s->src[k] = 0;
    const int id = s->comp[k];
    if (id) {
      const double rho = Q[0][k], g = s->Pg[k], cp = g * RGAS / (g - 1);
      const double T = s->Pp[k] / (rho * RGAS), Tt = T + 0.5 * (s->Pu[k] * s->Pu[k] + s->Pv[k] * s->Pv[k]) / cp;
      const double ty = cmp[id * 4 + 2], mf = mfx / ph;
      if (ty == 1 || ty == 2) {
        const double e = cmp[id * 4 + 1], fac = ty == 1 ? 1 : clampd(mf / (0.3 * (double)s->mfd[k]), 0, 1);
        const double mw = s->Fx[0][fx0], HU = mw > 1e-9 ? s->Fx[3][fx0] / mw : cp * Tt;
        const double w = (HU > 0 ? HU : cp * Tt) * expm1(cmp[id * 4] * (double)s->wgt[k] * fac * s->dx[i]) / s->dx[i];
        s->R[1][k] += ty == 1 ? e * rho * w : rho * w / e;
        s->R[3][k] += mf * w; s->src[k] = mf * w;
      } else if (s->rate[k] > 0) {
        const double K = cmp[id * 4 + 1] != 0 ? cmp[id * 4 + 1] : 1;
        const double tgt = cmp[id * 4] + ((double)s->tgt[k] - cmp[id * 4]) * K;
        const double m0 = jmax(mf, 0);
        const double q = m0 * cp * jmax(0, tgt - Tt) * (double)s->rate[k];
        s->R[3][k] += q;
        const double mfuel = q / s->hf;
        s->R[0][k] += mfuel; s->src[k] = mfuel;
        s->R[4][k] += m0 * (1 - s->PY[k]) * (double)s->yrate[k] + mfuel;
      }
    }
    if (computeDt) {
      const double c = sqrt(s->Pg[k] * s->Pp[k] / s->Pr[k]);
      s->dt[k] = s->cfl * jmax(s->phi[k], 0.25) /
                 ((fabs(s->Pu[k]) + c) / s->dx[i] + (fabs(s->Pv[k]) + c) / s->dr[j] + 2.2 * s->nu[k] / (s->dr[j] * s->dr[j]));
    }
  }
}

/* CPUSolver.fix: src/cfd_cpu.js:187-192 */
static void fixq(double *const *Q, int k) {
  const double r = jmax(Q[0][k], 1e-3);
  Q[0][k] = r;
  const double Y = clampd(Q[4][k] / r, 0, 1);
  Q[4][k] = Y * r;
  const double g = 1.4 - 0.07 * Y, ke = 0.5 * (Q[1][k] * Q[1][k] + Q[2][k] * Q[2][k]) / r;
  if ((g - 1) * (Q[3][k] - ke) < 50) Q[3][k] = 50 / (g - 1) + ke;
}

/* ---------------- C interface ---------------- */
static double *dup_d(const double *p, int n) { double *d = malloc(sizeof(double) * (size_t)n); memcpy(d, p, sizeof(double) * (size_t)n); return d; }
static float *dup_f(const float *p, int n) { float *d = malloc(sizeof(float) * (size_t)n); memcpy(d, p, sizeof(float) * (size_t)n); return d; }

AX *ax_create(int W, int H, const double *xf, const double *rf, const uint8_t *solid, const float *phi, const float *ae,
              const float *an, const uint8_t *comp, const float *wgt, const float *lmix, const float *tgt, const float *rate,
              const float *yrate, const float *mfd, const double *comps, const double *fs, const double *Q0) {
  AX *s = calloc(1, sizeof(AX));
  s->W = W; s->H = H; s->N = W * H;
  const int N = s->N;
  s->xc = malloc(8 * (size_t)W); s->dx = malloc(8 * (size_t)W); s->rc = malloc(8 * (size_t)H); s->dr = malloc(8 * (size_t)H);
  s->rf = dup_d(rf, H + 1);
  for (int i = 0; i < W; i++) { s->xc[i] = 0.5 * (xf[i] + xf[i + 1]); s->dx[i] = xf[i + 1] - xf[i]; }
  for (int j = 0; j < H; j++) { s->rc[j] = 0.5 * (rf[j] + rf[j + 1]); s->dr[j] = rf[j + 1] - rf[j]; }
  s->solid = malloc((size_t)N); memcpy(s->solid, solid, (size_t)N);
  s->comp = malloc((size_t)N); memcpy(s->comp, comp, (size_t)N);
  s->phi = dup_f(phi, N); s->ae = dup_f(ae, N); s->an = dup_f(an, N); s->wgt = dup_f(wgt, N); s->lmix = dup_f(lmix, N);
  s->tgt = dup_f(tgt, N); s->rate = dup_f(rate, N); s->yrate = dup_f(yrate, N); s->mfd = dup_f(mfd, N);
  memcpy(s->comps, comps, sizeof s->comps);
  s->fs_r = fs[0]; s->fs_u = fs[1]; s->fs_p = fs[2];
  s->cfl = 0.6; s->order = 2; s->visc = 1; s->hf = 0.99 * 43e6; s->global = 0; s->dtRatio = 0;
  for (int q = 0; q < 5; q++) {
    s->Q[q] = dup_d(Q0 + (size_t)q * N, N); s->Q1[q] = calloc((size_t)N, 8); s->R[q] = calloc((size_t)N, 8);
    s->Fx[q] = calloc((size_t)(W + 1) * H, 8); s->Fr[q] = calloc((size_t)W * (H + 1), 8);
  }
  s->dt = calloc((size_t)N, 8); s->nu = calloc((size_t)N, 8); s->src = calloc((size_t)N, 8); s->mfx = calloc((size_t)N, 8);
  s->Pr = calloc((size_t)N, 8); s->Pu = calloc((size_t)N, 8); s->Pv = calloc((size_t)N, 8); s->Pp = calloc((size_t)N, 8);
  s->PY = calloc((size_t)N, 8); s->Pg = calloc((size_t)N, 8);
  return s;
}
void ax_destroy(AX *s) {
  if (!s) return;
  void *p[] = {s->xc, s->rc, s->dx, s->dr, s->rf, s->solid, s->comp, s->phi, s->ae, s->an, s->wgt, s->lmix, s->tgt, s->rate,
               s->yrate, s->mfd, s->dt, s->nu, s->src, s->mfx, s->Pr, s->Pu, s->Pv, s->Pp, s->PY, s->Pg};
  for (size_t i = 0; i < sizeof p / sizeof p[0]; i++) free(p[i]);
  for (int q = 0; q < 5; q++) { free(s->Q[q]); free(s->Q1[q]); free(s->R[q]); free(s->Fx[q]); free(s->Fr[q]); }
  free(s);
}
/* opt (src/cfd_cpu.js:53, 63, 198-202): cfl, order, visc, hf (0 keeps 0.99 * 43e6), global, dtRatio (0 = off) */
void ax_set_opt(AX *s, double cfl, int order, int visc, double hf, int global, double dtRatio) {
  s->cfl = cfl; s->order = order; s->visc = visc; s->hf = hf > 0 ? hf : 0.99 * 43e6; s->global = global; s->dtRatio = dtRatio;
}
void ax_set_comps(AX *s, const double *c) { memcpy(s->comps, c, sizeof s->comps); }
/* CPUSolver.step: src/cfd_cpu.js:193-210 */
void ax_step(AX *s, int n) {
  const int N = s->N;
  const uint8_t *S = s->solid;
  for (int it = 0; it < n; it++) {
    residual(s, s->Q, 1);
    if (s->global || s->dtRatio != 0) {
      double mn = INFINITY;
      for (int k = 0; k < N; k++) if (!S[k] && s->dt[k] < mn) mn = s->dt[k];
      const double cap = s->global ? mn : mn * s->dtRatio;
      for (int k = 0; k < N; k++) if (s->dt[k] > cap) s->dt[k] = cap;
    }
    for (int q = 0; q < 5; q++) for (int k = 0; k < N; k++) s->Q1[q][k] = S[k] ? s->Q[q][k] : s->Q[q][k] + s->dt[k] * s->R[q][k];
    for (int k = 0; k < N; k++) if (!S[k]) fixq(s->Q1, k);
    residual(s, s->Q1, 0);
    for (int q = 0; q < 5; q++) for (int k = 0; k < N; k++) if (!S[k]) s->Q[q][k] = 0.5 * s->Q[q][k] + 0.5 * (s->Q1[q][k] + s->dt[k] * s->R[q][k]);
    for (int k = 0; k < N; k++) if (!S[k]) fixq(s->Q, k);
    s->iter++;
  }
}
void ax_read(const AX *s, double *Q, double *src, double *mfx, double *dt) {
  const size_t b = sizeof(double) * (size_t)s->N;
  for (int q = 0; q < 5; q++) memcpy(Q + (size_t)q * s->N, s->Q[q], b);
  if (src) memcpy(src, s->src, b);
  if (mfx) memcpy(mfx, s->mfx, b);
  if (dt) memcpy(dt, s->dt, b);
}
void ax_write(AX *s, const double *Q) { for (int q = 0; q < 5; q++) memcpy(s->Q[q], Q + (size_t)q * s->N, sizeof(double) * (size_t)s->N); }
