// ============================================================================
// field.hpp — the engine field model (FieldModel, src/field.js:255-488) and
// the axisymmetric solver's initial state (initialState with streamState1D,
// src/cfd_grid.js:242-257, 279-305), evaluated per point / per cell.
//
// The model arrives as one flat table of doubles (layout: AXI_FM_* in axi.h):
// stream walls, body polygons, the combustor / mixer / afterburner geometry
// and the constructor's products (stream knots with their Mach numbers,
// processed exits, inlet face velocity, hot gas). Everything here is FP64 in
// the reference's operation order with multiply-add contraction off, and
// JavaScript's max / min (NaN-propagating), so that comparisons that decide a
// region or a branch see the same numbers; pow / exp / sin / cos / tan / log
// are the device library's (within an ulp of a JS engine's).
//
// The functions are host + device so that a CPU build of this header can be
// checked against the reference's outputs before a device run.
// ============================================================================
#pragma once
#include <math.h>
#include <stdint.h>
#include "axi.h"

namespace fieldm {

// no multiply-add contraction: every product and sum rounded as written
#ifdef __clang__
#pragma clang fp contract(off)
#endif

#define FM_HD __host__ __device__ inline

constexpr double kPi = 3.141592653589793;
constexpr double kDeg = kPi / 180.0;
constexpr double kRGas = 287.05;                 // RGAS, src/cfd_grid.js:4

FM_HD double jmax(double a, double b) { return (a != a || b != b) ? (a != a ? a : b) : (a > b ? a : b); }
FM_HD double jmin(double a, double b) { return (a != a || b != b) ? (a != a ? a : b) : (a < b ? a : b); }
FM_HD double jclamp(double v, double a, double b) { return v < a ? a : v > b ? b : v; }
FM_HD double jlerp(double a, double b, double t) { return a + (b - a) * t; }
// smoothstep (src/field.js:6)
FM_HD double jsmooth(double e0, double e1, double x) {
  const double t = jclamp((x - e0) / (e1 - e0), 0.0, 1.0);
  return t * t * (3.0 - 2.0 * t);
}
FM_HD double jsign(double v) { return v > 0.0 ? 1.0 : v < 0.0 ? -1.0 : v; }
FM_HD int ii(double v) { return (int)v; }

// Polyline value at x (Line.at, src/geometry.js:27-34): flat beyond the ends,
// linear between the bracketing points found by bisection.
FM_HD double line_at(const double *T, double off, double npt, double x) {
  const double *p = T + ii(off);
  const int n = ii(npt);
  if (x <= p[0]) return p[1];
  if (x >= p[2 * (n - 1)]) return p[2 * (n - 1) + 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (p[2 * m] <= x) lo = m; else hi = m; }
  const double dx = p[2 * hi] - p[2 * lo];
  return dx < 1e-12 ? p[2 * hi + 1] : p[2 * lo + 1] + (p[2 * hi + 1] - p[2 * lo + 1]) * (x - p[2 * lo]) / dx;
}

// Even-odd crossing test (pointInPoly, src/geometry.js:61-68).
FM_HD bool in_poly(const double *T, double off, double npt, double x, double r) {
  const double *p = T + ii(off);
  const int n = ii(npt);
  bool inside = false;
  for (int i = 0, j = n - 1; i < n; j = i++) {
    const double xi = p[2 * i], ri = p[2 * i + 1], xj = p[2 * j], rj = p[2 * j + 1];
    if ((ri > r) != (rj > r) && x < (xj - xi) * (r - ri) / (rj - ri) + xi) inside = !inside;
  }
  return inside;
}

// Mass-flow parameter and its subsonic inverse by 48 bisections (mfp /
// invMfp, src/physics.js:29-42).
FM_HD double mfp(double M, double g, double R) {
  return M * sqrt(g / R) * pow(1.0 + 0.5 * (g - 1.0) * M * M, -(g + 1.0) / (2.0 * (g - 1.0)));
}
FM_HD double inv_mfp_sub(double val, double g, double R) {
  const double vmax = mfp(1.0, g, R);
  if (!(val > 0.0)) return 0.0;
  if (val >= vmax) return 1.0;
  double lo = 0.0, hi = 1.0;
  for (int i = 0; i < 48; i++) {
    const double mid = 0.5 * (lo + hi), v = mfp(mid, g, R);
    if (v < val) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

// Exponentially scaled I0 by the polynomial fits of Abramowitz & Stegun
// 9.8.1 / 9.8.2 (I0e, src/field.js:227-235).
FM_HD double i0e(double x) {
  const double ax = fabs(x);
  if (ax < 3.75) {
    const double u = x / 3.75, t = u * u;
    return exp(-ax) * (1.0 + t * (3.5156229 + t * (3.0899424 + t * (1.2067492 + t * (0.2659732 + t * (0.0360768 + t * 0.0045813))))));
  }
  const double t = 3.75 / ax;
  return (0.39894228 + t * (0.01328592 + t * (0.00225319 + t * (-0.00157565 + t * (0.00916281 + t * (-0.02057706 +
         t * (0.02635537 + t * (-0.01647633 + t * 0.00392377)))))))) / sqrt(ax);
}

// Share of a Gaussian-blurred disk of radius R at radius r (blurDisk,
// src/field.js:238-250): 10-point Gauss-Legendre over r +- 6 sigma.
FM_HD double blur_disk(double r, double R, double sig) {
  // 10-point Gauss-Legendre nodes and weights, to the reference's ten digits
  const double GX[10] = {-0.9739065285, -0.8650633667, -0.6794095683, -0.4333953941, -0.1488743390,
                         0.1488743390,  0.4333953941,  0.6794095683,  0.8650633667,  0.9739065285};
  const double GW[10] = {0.0666713443, 0.1494513492, 0.2190863625, 0.2692667193, 0.2955242247,
                         0.2955242247, 0.2692667193, 0.2190863625, 0.1494513492, 0.0666713443};
  if (sig < 1e-5) return r < R ? 1.0 : 0.0;
  const double a = jmax(0.0, r - 6.0 * sig), b = jmin(R, r + 6.0 * sig);
  if (b <= a) return 0.0;
  const double hm = 0.5 * (b - a), c = 0.5 * (a + b), is2 = 1.0 / (sig * sig);
  double sum = 0.0;
  for (int i = 0; i < 10; i++) {
    const double rho = c + hm * GX[i];
    const double d = r - rho;
    sum += GW[i] * rho * is2 * exp(-0.5 * d * d * is2) * i0e(r * rho * is2);
  }
  return jmin(1.0, sum * hm);
}

// Knot interpolation along a stream (knotAt, src/field.js:211-225): the two
// bracketing knots by bisection, linear in everything but Pt (geometric).
struct Knot { double Tt, Pt, M, sw, g, cp, R; };
FM_HD Knot knot_at(const double *T, const double *kn, int n, double x) {
  int lo = 0, hi = n - 1;
  if (x <= kn[0]) { lo = hi = 0; }
  else if (x >= kn[10 * hi]) { lo = hi; }
  else {
    while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (kn[10 * m] <= x) lo = m; else hi = m; }
  }
  const double *a = kn + 10 * lo, *b = kn + 10 * hi;
  const double t = (hi == lo || b[0] - a[0] < 1e-9) ? 0.0 : (x - a[0]) / (b[0] - a[0]);
  Knot o;
  o.Tt = jlerp(a[1], b[1], t); o.Pt = a[2] * pow(b[2] / a[2], t);
  o.M = jlerp(a[3], b[3], t); o.sw = jlerp(a[4], b[4], t);
  o.g = jlerp(a[5], b[5], t); o.cp = jlerp(a[6], b[6], t); o.R = jlerp(a[7], b[7], t);
  (void)T;
  return o;
}

// One sample of the model (the fields FieldModel.sample writes)
struct Sample {
  int region;        // 0 outside the engine, 1 solid, 2 + stream index
  double T, P, Tt, Pt, V, Vx, Vth, M, a, rho, g, cp, R;
};

// Region of (x, r) (classify, src/field.js:296-310): the first stream in the
// model's order whose walls bracket r at x, else the first body polygon that
// holds the point, else outside.
struct Cls { int kind, stream; double lo, hi; };   // kind 0 ext, 1 solid, 2 stream
FM_HD Cls classify(const double *T, double x, double r) {
  Cls c{0, -1, 0.0, 0.0};
  const int ns = ii(T[AXI_FM_NSTREAM]);
  for (int s = 0; s < ns; s++) {
    const double *S = T + AXI_FM_STREAM + 10 * s;
    if (x < S[1] || x > S[2]) continue;
    const double lo = line_at(T, S[3], S[4], x), hi = line_at(T, S[5], S[6], x);
    if (r >= lo && r <= hi) { c.kind = 2; c.stream = s; c.lo = lo; c.hi = hi; return c; }
  }
  const int nb = ii(T[AXI_FM_NBODY]);
  for (int b = 0; b < nb; b++) {
    const double *B = T + AXI_FM_BODY + 6 * b;
    if (x < B[0] || x > B[1] || r < B[2] || r > B[3]) continue;
    if (in_poly(T, B[4], B[5], x, r)) { c.kind = 1; return c; }
  }
  return c;
}

// Inside a stream (sampleStream, src/field.js:328-418): the knots' total
// state, with the combustor's temperature pattern, the mixer's lobe mixing
// and the afterburner's flame spread on top, static values from the Mach
// number and a 1/7-power boundary layer at the walls.
FM_HD Sample sample_stream(const double *T, double x, double r, double th, const Cls &cl) {
  const double *S = T + AXI_FM_STREAM + 10 * cl.stream;
  const Knot k = knot_at(T, T + ii(S[7]), ii(S[8]), x);
  const double *C = T + AXI_FM_CYCLE;          // Tt3, Tt4, Tt5, Tt6A, Tt7, Tt13, abOn
  const double s = T[AXI_FM_S];
  const double h = jmax(1e-6, cl.hi - cl.lo);
  const double psi = (r - cl.lo) / h;
  double Tt = k.Tt, M = k.M;
  const double g = k.g, cp = k.cp, R = k.R;
  // combustor
  const double *cb = T + AXI_FM_COMB;          // stream, x0, xd, x1, turbEnd, n, cannular, liners
  if (ii(cb[0]) == cl.stream && x >= cb[1] && x <= cb[4]) {
    const double Tt3 = C[0], Tt4 = C[1], dT = jmax(0.0, Tt4 - Tt3);
    const double N = cb[5], cosN = cos(N * th);
    const bool cann = cb[6] != 0.0;
    if (x <= cb[3]) {
      const double li = line_at(T, cb[7], cb[8], x), lo = line_at(T, cb[9], cb[10], x);
      if (x < cb[2] || r < li || r > lo) {
        Tt = Tt3 * (1.0 + 0.012 * jsmooth(cb[2], cb[3], x));
        M = 0.11;
      } else {
        const double sN = (x - cb[2]) / (cb[3] - cb[2]), pl = (r - li) / jmax(1e-6, lo - li);
        const double Tpz = jmin(jmin(Tt3 + 1.9 * dT, 2480.0), Tt3 + 1800.0);
        const double Tline = sN < 0.3 ? Tt3 + (Tpz - Tt3) * jsmooth(0.0, 0.26, sN)
                                      : Tt4 + (Tpz - Tt4) * exp(-3.4 * (sN - 0.3) / 0.7);
        const double amp = cann ? 0.55 : 0.33;
        const double q = (pl - 0.5) / 0.24;
        const double gP = (0.3 + 0.7 * exp(-(q * q))) * (1.0 + amp * cosN) / (1.0 + amp);
        const double gE = 1.0 + 0.32 * (sin(kPi * pl) - 2.0 / kPi) + (cann ? 0.14 : 0.08) * cosN;
        double G = jlerp(gP, gE, jsmooth(0.28, 1.0, sN));
        const double film = jsmooth(0.0, 0.07, pl) * jsmooth(0.0, 0.07, 1.0 - pl);
        G *= 0.35 + 0.65 * film;
        Tt = Tt3 + (Tline - Tt3) * G;
        M = jlerp(0.03, k.M, jsmooth(0.0, 0.7, sN));
      }
    } else {
      const double dec = exp(-(x - cb[3]) / (0.22 * s));
      const double pf = 0.32 * (sin(kPi * psi) - 2.0 / kPi) + (cann ? 0.14 : 0.08) * cosN;
      Tt = k.Tt + dT * pf * dec * 0.9;
    }
  }
  // mixer (the stream named ab)
  const double *mx = T + AXI_FM_MIXER;         // present, x, coreFrac, lobes
  if (mx[0] != 0.0 && ii(S[0]) == 3) {
    const double m = 1.0 - exp(-jmax(0.0, x - mx[1]) / (0.5 * s));
    const double lo = cl.lo, hi = cl.hi;
    const double pa = (r * r - lo * lo) / jmax(1e-9, hi * hi - lo * lo);
    const double pc = mx[2] + 0.13 * cos(mx[3] * th) * (1.0 - m);
    const double w = 0.03 + 0.5 * m;
    const double chi = 1.0 - jsmooth(pc - w, pc + w, pa);
    const double Tloc = chi * C[2] + (1.0 - chi) * C[5];
    Tt = Tt + (1.0 - m) * (Tloc - C[3]);
  }
  // afterburner
  const double *ab = T + AXI_FM_AB;            // present, stream, xFH, x1, rings (offset, count)
  if (ab[0] != 0.0 && ii(ab[1]) == cl.stream && x > ab[2] && C[6] != 0.0) {
    const double Tt6 = ii(T[AXI_FM_TYPE]) == 2 ? C[3] : C[2];
    const double L = ab[3] - ab[2], t = (x - ab[2]) / L;
    double F = 0.0;
    const double spread = (x - ab[2]) * tan(7.0 * kDeg) + 0.004 * s;
    const double *rg = T + ii(ab[4]);
    for (int q = 0; q < ii(ab[5]); q++) F = jmax(F, 1.0 - jsmooth(spread * 0.55, spread, fabs(r - rg[q])));
    F = jmax(F, jsmooth(0.45, 1.0, t));
    if (x >= ab[3]) F = 1.0;
    F *= jsmooth(0.0, 0.035 * s, cl.hi - r) * 0.85 + 0.15;
    const double meanT = k.Tt;
    const double TtF = Tt6 + (C[4] - Tt6) * F;
    Tt = x >= ab[3] ? meanT + (TtF - C[4]) : TtF + (Tt - Tt6) * (1.0 - F);
  }
  // static values
  const double kk = 1.0 + 0.5 * (g - 1.0) * M * M;
  const double P = k.Pt / pow(1.0 + 0.5 * (g - 1.0) * k.M * k.M, g / (g - 1.0));
  double Ts = Tt / kk;
  double V = M * sqrt(g * R * Ts);
  // boundary layer
  const double d = jmin(psi, 1.0 - psi) * h;
  const double del = jmax(0.07 * h, 0.004 * s);
  if (d < del) {
    const double u = pow(jmax(0.0, d) / del, 1.0 / 7.0);
    V *= u;
    Ts = Tt - V * V / (2.0 * cp);
  }
  const double a = sqrt(g * R * Ts);
  const double sw = k.sw * kDeg;
  Sample o;
  o.region = 2 + cl.stream;
  o.T = Ts; o.P = P; o.Tt = Tt; o.Pt = P * pow(Tt / Ts, g / (g - 1.0));
  o.V = V; o.Vx = V * cos(sw); o.Vth = V * sin(sw); o.M = V / a; o.a = a;
  o.rho = P / (R * Ts); o.g = g; o.cp = cp; o.R = R;
  return o;
}

// Outside the engine (sampleExt, src/field.js:420-488): freestream with the
// inlet sink and the nacelle-shoulder speed-up, the exits' jets as blurred
// disks carrying momentum and enthalpy fluxes, shock cells of under-expanded
// jets, hot-gas properties blended by the total-temperature rise.
FM_HD Sample sample_ext(const double *T, double x, double r) {
  const double *fl = T + AXI_FM_FL;            // V0, T0, P0, Tt0, Pt0
  const double *lip = T + AXI_FM_LIP;          // x, r, rOut, outer (offset, points), outerEnd
  const double *hot = T + AXI_FM_HOT;          // Uface, maxdT, hot present, g, cp
  const double g = T[AXI_FM_COLD], cp = T[AXI_FM_COLD + 1], R = cp * (g - 1.0) / g;
  const double U0 = fl[0], rl = lip[1];
  double U = U0;
  {
    const double dx = x <= lip[0] ? lip[0] - x : (x - lip[0]) * 4.0;
    const double dr = jmax(0.0, r - rl);
    const double q = dr / rl;
    const double f = (1.0 - dx / sqrt(dx * dx + rl * rl)) * exp(-1.6 * (q * q)) * (r < lip[2] * 1.02 || x < lip[0] ? 1.0 : 0.4);
    U += (hot[0] - U0) * f;
  }
  if (U0 > 1.0 && x > lip[0] && x < lip[5]) {
    const double L = lip[5] - lip[0], dn = jmax(0.0, r - line_at(T, lip[3], lip[4], x));
    const double sn = sin(kPi * (x - lip[0]) / L);
    U += 0.1 * U0 * (sn * sn) * exp(-dn / (0.5 * rl));
  }
  double Tt = fl[3];
  int zone = 0;                                // 0 free, 1 jet, 2 shock
  double qm = 0.0, qh = 0.0, shock = 0.0;
  int wcore = -1;
  const int ne = ii(T[AXI_FM_NEXIT]);
  for (int q = 0; q < ne; q++) {
    const double *e = T + AXI_FM_EXIT + 12 * q;   // x, R, spread, Qm, Qh, under, Ls, P, Pt, g, cp, dTt
    if (x < e[0]) continue;
    const double dxe = x - e[0];
    const double sig = 0.015 * e[1] + e[2] * dxe;
    const double w = blur_disk(r, e[1], sig);
    if (w <= 1e-6) continue;
    qm += w * e[3]; qh += w * e[4];
    if (e[5] != 0.0 && e[6] > 0.0) {
      const double env = exp(-dxe / (3.2 * e[6])) * pow(w, 6.0);
      if (env > 1e-3) { shock = env; wcore = q; }
    }
  }
  if (qm != 0.0 || qh != 0.0) {
    const double Uex = jsign(qm) * sqrt(fabs(qm));
    const double sU = fabs(Uex) < 2.0 ? 2.0 * jsign((Uex != 0.0 && Uex == Uex) ? Uex : 1.0) : Uex;
    double dTt = qh / sU;
    dTt = jclamp(dTt, -hot[1], hot[1]);
    U = U0 + Uex; Tt = fl[3] + dTt;
    zone = 1;
  }
  double Ts = Tt - U * U / (2.0 * cp);
  double P = fl[2];
  if (zone == 0) P = fl[4] * pow(jmax(Ts, 1.0) / fl[3], g / (g - 1.0));
  if (shock > 0.0) {
    const double *e = T + AXI_FM_EXIT + 12 * wcore;
    const double dxe = x - e[0];
    const double Pc = fl[2] * exp(log(e[7] / fl[2]) * cos(2.0 * kPi * dxe / e[6]) * exp(-dxe / (3.2 * e[6])));
    const double ge = e[9], cpe = e[10];
    const double Tiso = Tt * pow(jmin(1.0, Pc / e[8]), (ge - 1.0) / ge);
    const double Uiso = sqrt(jmax(0.0, 2.0 * cpe * (Tt - Tiso)));
    const double b = jmin(1.0, shock);
    P = jlerp(P, Pc, b); U = jlerp(U, Uiso, b); Ts = Tt - U * U / (2.0 * cp);
    zone = 2;
  }
  Ts = jmax(Ts, 50.0);
  double gg = g, cpp = cp, RR = R;
  if (zone != 0 && hot[2] != 0.0) {
    const double fr = jclamp((Tt - fl[3]) / hot[1], 0.0, 1.0);
    const double hg = hot[3], hcp = hot[4];
    cpp = jlerp(cp, hcp, fr); RR = jlerp(R, hcp * (hg - 1.0) / hg, fr); gg = cpp / (cpp - RR);
    Ts = jmax(50.0, Tt - U * U / (2.0 * cpp));
  }
  const double a = sqrt(gg * RR * Ts);
  Sample o;
  o.region = 0;
  o.T = Ts; o.P = P; o.Tt = Tt; o.Pt = P * pow(Tt / Ts, gg / (gg - 1.0));
  o.V = fabs(U); o.Vx = U; o.Vth = 0.0; o.M = fabs(U) / a; o.a = a;
  o.rho = P / (RR * Ts); o.g = gg; o.cp = cpp; o.R = RR;
  return o;
}

// FieldModel.prototype.sample (src/field.js:318-326)
FM_HD Sample sample(const double *T, double x, double r, double th, Cls *clout = nullptr) {
  const Cls cl = classify(T, x, r);
  if (clout) *clout = cl;
  if (cl.kind == 1) {
    Sample o;
    const double nan = __builtin_nan("");
    o.region = 1;
    o.T = o.P = o.Tt = o.Pt = o.V = o.Vx = o.Vth = o.M = o.a = o.rho = o.g = o.cp = o.R = nan;
    return o;
  }
  if (cl.kind == 2) return sample_stream(T, x, r, th, cl);
  return sample_ext(T, x, r);
}

// One-dimensional state of stream s at x (streamState1D, src/cfd_grid.js:242-257):
// knot interpolation, the Mach number from the mass flow through the stream's
// annulus (or the nozzle knots' own), capped at 0.95.
struct State1D { double T, p, u, Tt; };
FM_HD State1D stream_state_1d(const double *T, int s, double x) {
  const double *S = T + AXI_FM_STREAM + 10 * s;
  const double *kn = T + ii(S[7]);
  const int n = ii(S[8]);
  int lo = 0, hi = n - 1;
  if (x <= kn[0]) hi = 0;
  else if (x >= kn[10 * hi]) lo = hi;
  else while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (kn[10 * m] <= x) lo = m; else hi = m; }
  const double *a = kn + 10 * lo, *b = kn + 10 * hi;
  const double t = (hi == lo || b[0] - a[0] < 1e-9) ? 0.0 : (x - a[0]) / (b[0] - a[0]);
  const double Tt = jlerp(a[1], b[1], t), Pt = a[2] * pow(b[2] / a[2], t), g = jlerp(a[5], b[5], t), R = jlerp(a[7], b[7], t);
  const double md = jlerp(a[8], b[8], t);
  const double rh = line_at(T, S[5], S[6], x), rl = line_at(T, S[3], S[4], x);
  const double A = kPi * (rh * rh - rl * rl);
  double M;
  if (a[9] >= 0.0 && b[9] >= 0.0) M = jlerp(a[3], b[3], t);
  else M = inv_mfp_sub(md * sqrt(Tt) / (Pt * jmax(A, 1e-6)), g, R);
  M = jmin(M, 0.95);
  const double k = 1.0 + 0.5 * (g - 1.0) * M * M, Ts = Tt / k, p = Pt / pow(k, g / (g - 1.0));
  State1D o;
  o.T = Ts; o.p = p; o.u = M * sqrt(g * R * Ts); o.Tt = Tt;
  return o;
}

// initialState of cell (i, j) (src/cfd_grid.js:279-305) at centre (x, r):
// rest at the freestream's static state in solids, else the model's sample
// with the one-dimensional state of the cell's own stream, and the burnt-gas
// fraction Y behind the combustor, in the afterburner stream and in a hot jet.
struct Cons { double q[5]; };
FM_HD Cons initial_cell(const double *T, double x, double r, int solid, int stream, int uniform) {
  const double *fl = T + AXI_FM_FL;            // V0, T0, P0, Tt0, Pt0
  const double *C = T + AXI_FM_CYCLE;
  double rho, u, p, Y = 0.0;
  if (uniform || solid) {
    p = fl[2]; rho = fl[2] / (kRGas * fl[1]); u = uniform ? fl[0] : 0.0;
  } else {
    Sample o = sample(T, x, r, 0.0);
    if (o.region == 1) {
      p = fl[2]; rho = fl[2] / (kRGas * fl[1]); u = 0.0;
    } else {
      const int st = stream;                   // index into the model's stream order, < 0: none
      if (st >= 0) { const State1D q = stream_state_1d(T, st, x); o.P = q.p; o.T = q.T; o.Vx = q.u; o.Tt = q.Tt; }
      p = o.P; rho = o.P / (kRGas * o.T); u = o.Vx;
      const double *cb = T + AXI_FM_COMB;
      if (st >= 0 && st == ii(cb[0]) && x > cb[2]) Y = jsmooth(cb[2], cb[3], x);
      if (st >= 0 && ii(T[AXI_FM_STREAM + 10 * st]) == 3) Y = jclamp((o.Tt - C[5]) / jmax(1.0, C[2] - C[5]), 0.0, 1.0);
      if (st < 0) {
        // the exit with the largest total-temperature rise (the first of equals)
        int he = -1;
        double best = -1e9;
        for (int q = 0; q < ii(T[AXI_FM_NEXIT]); q++) {
          const double d = T[AXI_FM_EXIT + 12 * q + 11];
          if (d > best) { best = d; he = q; }
        }
        if (he >= 0) {
          const double *e = T + AXI_FM_EXIT + 12 * he;
          if (e[11] > 50.0 && x > e[0]) Y = jclamp((o.Tt - fl[3]) / e[11], 0.0, 1.0);
        }
      }
    }
  }
  const double g = 1.4 - 0.07 * Y;             // gamOfY, src/cfd_grid.js:5
  Cons c;
  c.q[0] = rho; c.q[1] = rho * u; c.q[2] = 0.0; c.q[3] = p / (g - 1.0) + 0.5 * rho * u * u; c.q[4] = rho * Y;
  return c;
}

// Host-side check of a table before any device use: version, counts within
// their caps, every offset and extent inside the array, every knot list and
// polyline non-empty. Returns null when the table is sound, else the reason.
inline const char *check_table(const double *T, int n) {
  if (!T || n < AXI_FM_HEAD) return "field table: shorter than its header";
  auto isint = [](double v) { return v == v && v >= 0.0 && v <= 1e9 && (double)(long long)v == v; };
  auto span = [&](double off, double cnt, int width, int minc) -> bool {
    if (!isint(off) || !isint(cnt) || cnt < minc) return false;
    const long long o = (long long)off, c = (long long)cnt;
    return o >= AXI_FM_HEAD && o + c * width <= n;
  };
  if (T[AXI_FM_VERSION] != 1.0) return "field table: version is not 1";
  if (!(T[AXI_FM_TYPE] == 0.0 || T[AXI_FM_TYPE] == 1.0 || T[AXI_FM_TYPE] == 2.0)) return "field table: engine type not 0, 1 or 2";
  const double ns = T[AXI_FM_NSTREAM], nb = T[AXI_FM_NBODY], ne = T[AXI_FM_NEXIT];
  if (!isint(ns) || ns < 1 || ns > 4) return "field table: stream count not 1..4";
  if (!isint(nb) || nb > 16) return "field table: body count not 0..16";
  if (!isint(ne) || ne > 4) return "field table: exit count not 0..4";
  for (int s = 0; s < (int)ns; s++) {
    const double *S = T + AXI_FM_STREAM + 10 * s;
    if (!isint(S[0]) || S[0] > 3) return "field table: stream name code not 0..3";
    if (!span(S[3], S[4], 2, 1) || !span(S[5], S[6], 2, 1)) return "field table: stream wall outside the table";
    if (!span(S[7], S[8], 10, 1)) return "field table: stream knots outside the table";
  }
  for (int b = 0; b < (int)nb; b++)
    if (!span(T[AXI_FM_BODY + 6 * b + 4], T[AXI_FM_BODY + 6 * b + 5], 2, 1)) return "field table: body polygon outside the table";
  const double *cb = T + AXI_FM_COMB;
  if (!isint(cb[0]) || cb[0] >= ns) return "field table: combustor stream index out of range";
  if (!span(cb[7], cb[8], 2, 1) || !span(cb[9], cb[10], 2, 1)) return "field table: combustor liner outside the table";
  const double *ab = T + AXI_FM_AB;
  if (ab[0] != 0.0) {
    if (!isint(ab[1]) || ab[1] >= ns) return "field table: afterburner stream index out of range";
    if (!span(ab[4], ab[5], 1, 0)) return "field table: afterburner rings outside the table";
  }
  if (!span(T[AXI_FM_LIP + 3], T[AXI_FM_LIP + 4], 2, 1)) return "field table: outer line outside the table";
  return nullptr;
}

#undef FM_HD

}  // namespace fieldm
