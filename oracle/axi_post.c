/* ============================================================================
 * axi_post.c — CPU restatement of the axisymmetric post-processing the
 * reference runs on a solver's state (the device versions are checked
 * against it):
 *   oax_thrust   cfdForces with the solver's own sources (src/cfd_post.js:44-75)
 *   oax_section  the sums of cfdSection (src/cfd_post.js:83-102), which also
 *                give the station averages of tools/stations.js:8-16
 *   oax_derived  cfdDerived's display fields (src/cfd_post.js:10-40)
 *   oax_interp   interpState, grid sequencing (src/app_cfd.js:321-358)
 * TEST INFRASTRUCTURE ONLY: the checker of libaero3d.so; the product never
 * links it. Each function is the reference's double arithmetic in its loop
 * order (built with -ffp-contract=off), so on the reference's own states it
 * reproduces the reference's outputs (pinned on its fixtures in the tests).
 * Grids: xc, dx (W) and rc, dr (H) in doubles as buildCFDGrid forms them
 * (centre = half the sum of the faces, width = difference); states: 5 planes
 * of N = W H doubles, row j and column i at j W + i.
 * ========================================================================== */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define RGAS 287.05                                   /* src/cfd_grid.js:4 */

static inline double jsmax(double a, double b) { return (a != a || b != b) ? (a != a ? a : b) : (a > b ? a : b); }
static inline double jsmin(double a, double b) { return (a != a || b != b) ? (a != a ? a : b) : (a < b ? a : b); }
static inline double clampv(double v, double a, double b) { return v < a ? a : v > b ? b : v; }

typedef struct { double r, u, v, p, Y, g; } Prim;
/* cfdPrim (src/cfd_post.js:4-8) */
static Prim prim_of(const double *Q, int N, int k) {
  Prim q;
  q.r = jsmax(Q[k], 1e-4);
  q.u = Q[N + k] / q.r;
  q.v = Q[2 * N + k] / q.r;
  q.Y = clampv(Q[4 * N + k] / q.r, 0, 1);
  q.g = 1.4 - 0.07 * q.Y;
  q.p = jsmax(1, (q.g - 1) * (Q[3 * N + k] - 0.5 * q.r * (q.u * q.u + q.v * q.v)));
  return q;
}

/* out: Fb, Fw, mfuel, pw[0..7] */
void oax_thrust(int W, int H, const double *rc, const double *dr, const double *dx, const uint8_t *solid, const float *phi,
                const float *ae, const uint8_t *comp, const float *rate, const double *Q, const double *S, const double *F,
                const double *comps, double P0, double *out) {
  const int N = W * H;
  double Fb = 0, Fw = 0, mfuel = 0, pw[8] = {0};
  for (int j = 0; j < H; j++)
    for (int i = 0; i < W; i++) {
      const int k = j * W + i;
      if (solid[k]) continue;
      const double A = 2 * M_PI * rc[j] * dr[j], V = A * dx[i] * phi[k];
      const int id = comp[k];
      if (id) {
        const double ty = comps[id * 4 + 2];
        if (ty == 1 || ty == 2) {
          const Prim q = prim_of(Q, N, k);
          const double e = comps[id * 4 + 1], mf = F[k] / phi[k];
          const double w = fabs(mf) > 1e-9 ? S[k] / mf : 0;
          Fb += (ty == 1 ? e * q.r * w : q.r * w / e) * V;
          pw[id] += S[k] * V;
        } else if (rate[k] > 0) {
          mfuel += S[k] * V;
        }
      }
      const double aw = i > 0 ? ae[k - 1] : 1, dA = ae[k] - aw;
      if (dA != 0) Fw += (prim_of(Q, N, k).p - P0) * dA * A;
    }
  out[0] = Fb; out[1] = Fw; out[2] = mfuel;
  memcpy(out + 3, pw, sizeof pw);
}

/* sums: mass flow, w, A, sum p dA, then the w-weighted Tt, Pt, T, p, |V|, M, Y */
void oax_section(int W, int H, const double *rc, const double *dr, const float *ae, const double *Q, const double *mfx, int col,
                 const uint8_t *rows, double *sums) {
  const int N = W * H, i = col;
  double a[11] = {0};
  for (int j = 0; j < H; j++) {
    if (!rows[j]) continue;
    const int k = j * W + i;
    const Prim q = prim_of(Q, N, k);
    const double al = jsmin(1, 0.5 * ((i > 0 ? (double)ae[k - 1] : 1.0) + (double)ae[k])), dA = 2 * M_PI * rc[j] * dr[j] * al;
    const double T = q.p / (q.r * RGAS), cp = q.g * RGAS / (q.g - 1), V2 = q.u * q.u + q.v * q.v, Tt = T + V2 / (2 * cp);
    const double Pt = q.p * pow(Tt / T, q.g / (q.g - 1)), mf = mfx[k] * 2 * M_PI * rc[j] * dr[j], wk = jsmax(mf, 0);
    a[0] += mf; a[1] += wk; a[2] += dA; a[3] += q.p * dA;
    a[4] += wk * Tt; a[5] += wk * Pt; a[6] += wk * T; a[7] += wk * q.p; a[8] += wk * sqrt(V2);
    a[9] += wk * sqrt(V2 / (q.g * q.p / q.r)); a[10] += wk * q.Y;
  }
  memcpy(sums, a, sizeof a);
}

/* fields: 13 planes (T, Tt, P, Pt, V, M, rho, Vx, Vth, Vr, Y, vort, ds) */
void oax_derived(int W, int H, const double *xc, const double *rc, const uint8_t *solid, const float *phi, const int8_t *stream,
                 const uint8_t *mat, const double *Q, double T0, double P0, int disp, float *fields, uint8_t *reg, uint8_t *matout) {
  const int N = W * H;
  float *f[13];
  for (int c = 0; c < 13; c++) f[c] = fields + (size_t)c * N;
#define SOL(c) (solid[c] || (disp && phi[c] < 0.5))
  for (int k = 0; k < N; k++) {
    matout[k] = 0;
    f[8][k] = 0;
    if (SOL(k)) {
      reg[k] = 1; matout[k] = mat[k];
      for (int c = 0; c < 13; c++) if (c != 8) f[c][k] = NAN;
      continue;
    }
    reg[k] = stream[k] < 0 ? 0 : (uint8_t)(2 + stream[k]);
    const Prim q = prim_of(Q, N, k);
    const double T = q.p / (q.r * RGAS), cp = q.g * RGAS / (q.g - 1), V2 = q.u * q.u + q.v * q.v, a = sqrt(q.g * q.p / q.r);
    const double Tt = T + V2 / (2 * cp);
    f[0][k] = (float)T; f[2][k] = (float)q.p; f[6][k] = (float)q.r; f[7][k] = (float)q.u; f[9][k] = (float)q.v;
    f[4][k] = (float)sqrt(V2); f[5][k] = (float)(sqrt(V2) / a);
    f[1][k] = (float)Tt; f[3][k] = (float)(q.p * pow(Tt / T, q.g / (q.g - 1))); f[10][k] = (float)q.Y;
    f[12][k] = (float)(cp * log(T / T0) - RGAS * log(q.p / P0));
  }
  for (int j = 0; j < H; j++)
    for (int i = 0; i < W; i++) {
      const int k = j * W + i;
      if (SOL(k)) continue;
      const int kl = i > 0 && !SOL(k - 1) ? k - 1 : k, kr = i < W - 1 && !SOL(k + 1) ? k + 1 : k;
      const int kd = j > 0 && !SOL(k - W) ? k - W : k, ku = j < H - 1 && !SOL(k + W) ? k + W : k;
      const double dxx = xc[kr % W] - xc[kl % W], drr = rc[ku / W] - rc[kd / W];
      const double dvdx = dxx > 0 ? ((double)f[9][kr] - (double)f[9][kl]) / dxx : 0;
      const double dudr = drr > 0 ? ((double)f[7][ku] - (double)f[7][kd]) / drr : 0;
      f[11][k] = (float)(dudr - dvdx);
    }
#undef SOL
}

/* fracIdx (src/app_cfd.js:321-326) */
static double frac_idx(const double *a, int n, double v) {
  if (v <= a[0]) return (v - a[0]) / (a[1] - a[0]);
  if (v >= a[n - 1]) return n - 1 + (v - a[n - 1]) / (a[n - 1] - a[n - 2]);
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (a[m] <= v) lo = m; else hi = m; }
  return lo + (v - a[lo]) / (a[hi] - a[lo]);
}
static inline int clampi(int v, int a, int b) { return v < a ? a : v > b ? b : v; }

void oax_interp(int WA, int HA, const double *xcA, const double *rcA, const uint8_t *solidA, const double *QA, int WB, int HB,
                const double *xcB, const double *rcB, double T0, double P0, double *QB) {
  const int NA = WA * HA, NB = WB * HB;
  for (int j = 0; j < HB; j++) {
    const double fj = frac_idx(rcA, HA, rcB[j]);
    for (int i = 0; i < WB; i++) {
      const double fi = frac_idx(xcA, WA, xcB[i]);
      const int i0 = clampi((int)floor(fi), 0, WA - 1), j0 = clampi((int)floor(fj), 0, HA - 1);
      const double tx = clampv(fi - i0, 0, 1), ty = clampv(fj - j0, 0, 1);
      const int di_[4] = {0, 1, 0, 1}, dj_[4] = {0, 0, 1, 1};
      const double w_[4] = {(1 - tx) * (1 - ty), tx * (1 - ty), (1 - tx) * ty, tx * ty};
      double acc[5] = {0}, wsum = 0;
      for (int c = 0; c < 4; c++) {
        const int ii = i0 + di_[c] < WA - 1 ? i0 + di_[c] : WA - 1, jj = j0 + dj_[c] < HA - 1 ? j0 + dj_[c] : HA - 1, k = jj * WA + ii;
        if (solidA[k] || w_[c] <= 0) continue;
        for (int q = 0; q < 5; q++) acc[q] += w_[c] * QA[(size_t)q * NA + k];
        wsum += w_[c];
      }
      const int kB = j * WB + i;
      if (wsum > 0) { for (int q = 0; q < 5; q++) QB[(size_t)q * NB + kB] = acc[q] / wsum; continue; }
      int best = -1, bd = 1000000000;
      for (int dj = -3; dj <= 3; dj++)
        for (int di = -3; di <= 3; di++) {
          const int ii = clampi(i0 + di, 0, WA - 1), jj = clampi(j0 + dj, 0, HA - 1), k = jj * WA + ii;
          if (!solidA[k] && di * di + dj * dj < bd) { bd = di * di + dj * dj; best = k; }
        }
      for (int q = 0; q < 5; q++)
        QB[(size_t)q * NB + kB] = best >= 0 ? QA[(size_t)q * NA + best] : q == 0 ? P0 / (RGAS * T0) : q == 3 ? P0 / 0.4 : 0;
    }
  }
}
