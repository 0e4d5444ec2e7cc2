'use strict';
/* ============================================================================
 * cfd_post.js — post-processing of the axisymmetric engine solver on the host
 * (drop-in for src/cfd_post.js:4-126 of the reference): primitives of one cell,
 * thrust (blade body force + wall gauge pressure), mass-weighted section
 * averages, the closed-loop combustor / afterburner exit-temperature control,
 * spool power ratios. Each function keeps the operation order of the
 * reference expression it restates (its outputs are compared bit for bit).
 * ========================================================================== */
const RGAS = 287.05;                                               // src/cfd_grid.js:4
const clamp = (v, a, b) => (v < a ? a : v > b ? b : v);           // src/field.js:4

/* Primitives of cell k (src/cfd_post.js:4-8). */
function cfdPrim(Q, k) {
  const r = Math.max(Q[0][k], 1e-4), u = Q[1][k] / r, v = Q[2][k] / r, Y = clamp(Q[4][k] / r, 0, 1), g = 1.4 - 0.07 * Y;
  const p = Math.max(1, (g - 1) * (Q[3][k] - 0.5 * r * (u * u + v * v)));
  return { r, u, v, p, Y, g };
}

/* Axial force of the engine on the fluid (src/cfd_post.js:44-75); src =
 * {S: solver source per fluid volume, F: mean axial face mass flux} when the
 * solver's own sources are used. */
function cfdForces(grid, zones, Q, comps, P0, hf, src) {
  const { W, H, rc, dx, dr } = grid, S = zones.solid;
  let Fb = 0, Fw = 0, mfuel = 0;
  const pw = new Float64Array(8);
  for (let j = 0; j < H; j++) for (let i = 0; i < W; i++) {
    const k = j * W + i;
    if (S[k]) continue;
    const A = 2 * Math.PI * rc[j] * dr[j], V = A * dx[i] * zones.phi[k];
    const id = zones.comp[k];
    if (id) {
      const ty = comps[id * 4 + 2];
      if (ty === 1 || ty === 2) {
        const q = cfdPrim(Q, k), cp = q.g * RGAS / (q.g - 1), T = q.p / (q.r * RGAS), Tt = T + 0.5 * (q.u * q.u + q.v * q.v) / cp;
        const e = comps[id * 4 + 1], mf = src ? src.F[k] / zones.phi[k] : Q[1][k];
        const fac = ty === 1 ? 1 : clamp(mf / (0.3 * zones.mfd[k]), 0, 1);
        const w = src ? (Math.abs(mf) > 1e-9 ? src.S[k] / mf : 0) : cp * Tt * Math.expm1(comps[id * 4] * zones.wgt[k] * fac * dx[i]) / dx[i];
        Fb += (ty === 1 ? e * q.r * w : q.r * w / e) * V;
        pw[id] += (src ? src.S[k] : mf * w) * V;
      } else if (zones.rate[k] > 0) {
        if (src) mfuel += src.S[k] * V;
        else {
          const q = cfdPrim(Q, k), cp = q.g * RGAS / (q.g - 1), T = q.p / (q.r * RGAS), Tt = T + 0.5 * (q.u * q.u + q.v * q.v) / cp;
          const tgt = comps[id * 4] + (zones.tgt[k] - comps[id * 4]) * (comps[id * 4 + 1] || 1);
          mfuel += Math.max(Q[1][k], 0) * cp * Math.max(0, tgt - Tt) * zones.rate[k] / (hf || 0.99 * 43e6) * V;
        }
      }
    }
    const aw = i > 0 ? zones.ae[k - 1] : 1, dA = zones.ae[k] - aw;
    if (dA !== 0) Fw += (cfdPrim(Q, k).p - P0) * dA * A;
  }
  return { F: Fb + Fw, Fb, Fw, pw, mfuel };
}

/* Column nearest to x (src/cfd_post.js:76-80). */
function colAt(grid, x) {
  const { xc, W } = grid;
  let lo = 0, hi = W - 1;
  while (hi - lo > 1) { const m = (lo + hi) >> 1; if (xc[m] <= x) lo = m; else hi = m; }
  return Math.abs(xc[lo] - x) < Math.abs(xc[hi] - x) ? lo : hi;
}

/* Mass-weighted section averages at x over stream si (null: all fluid)
 * (src/cfd_post.js:83-102). */
function cfdSection(grid, zones, Q, x, si, mfx) {
  const { W, H, rc, dr } = grid, S = zones.solid;
  const i = colAt(grid, x);
  let m = 0, w = 0, A = 0, pA = 0;
  const acc = { Tt: 0, Pt: 0, T: 0, P: 0, V: 0, M: 0, Y: 0 };
  for (let j = 0; j < H; j++) {
    const k = j * W + i;
    if (S[k]) continue;
    if (si !== null && zones.stream[k] !== si) continue;
    const q = cfdPrim(Q, k), al = Math.min(1, 0.5 * ((i > 0 ? zones.ae[k - 1] : 1) + zones.ae[k])), dA = 2 * Math.PI * rc[j] * dr[j] * al;
    const T = q.p / (q.r * RGAS), cp = q.g * RGAS / (q.g - 1), V2 = q.u * q.u + q.v * q.v, Tt = T + V2 / (2 * cp);
    const Pt = q.p * Math.pow(Tt / T, q.g / (q.g - 1)), mf = mfx ? mfx[k] * 2 * Math.PI * rc[j] * dr[j] : q.r * q.u * dA, wk = Math.max(mf, 0);
    m += mf; w += wk; A += dA; pA += q.p * dA;
    acc.Tt += wk * Tt; acc.Pt += wk * Pt; acc.T += wk * T; acc.P += wk * q.p; acc.V += wk * Math.sqrt(V2);
    acc.M += wk * Math.sqrt(V2 / (q.g * q.p / q.r)); acc.Y += wk * q.Y;
  }
  const d = w || 1, out = { x: grid.xc[i], mdot: m, A };
  for (const key in acc) out[key] = acc[key] / d;
  if (!w) out.P = pA / (A || 1);
  return out;
}

const SPOOLS = { hbtf: [[[1, 2], 6], [[3], 5]], tj: [[[3], 5]], lbtf: [[[1], 6], [[3], 5]] };   // src/cfd_post.js:106
/* Turbine-to-compressor power ratio per spool (src/cfd_post.js:124-126). */
function spoolRatios(type, pw, etam) {
  return SPOOLS[type].map(([cs, t]) => { const Pc = cs.reduce((a, id) => a + pw[id], 0), Pt = -pw[t]; return Pc > 0 && Pt > 0 ? etam * Pt / Pc : NaN; });
}

/* Exit-temperature control of the heat sources (src/cfd_post.js:110-123): the
 * gain K of comps[id*4+1] moves toward target Tt, limited to 3 % per 100
 * iterations. ctl = {s, comb: {x1, stream}, ab: {x1, stream} | null, order,
 * Tt4, Tt7, abOn} — the spec / operating-point fields the reference reads. */
function cfdControl(type, ctl, grid, zones, Q, mfx, comps, pw, etam, dIt) {
  return controlFromSections(type, ctl, st => cfdSection(grid, zones, Q, st.x, st.si, mfx), comps, pw, etam, dIt);
}

/* The control stations of cfdControl: heat source id, section x and stream,
 * target exit temperature (combustor; afterburner when lit). */
function controlStations(ctl) {
  const out = [{ id: 4, x: ctl.comb.x1 - 0.004 * ctl.s, si: ctl.order.indexOf(ctl.comb.stream), target: ctl.Tt4 }];
  if (ctl.ab && ctl.abOn) out.push({ id: 7, x: ctl.ab.x1 - 0.004 * ctl.s, si: ctl.order.indexOf(ctl.ab.stream), target: ctl.Tt7 });
  return out;
}

/* cfdControl's gain update with the section of each station from secOf(st)
 * ({mdot, Tt}: cfdSection on the host, or the device monitor's sums). */
function controlFromSections(type, ctl, secOf, comps, pw, etam, dIt) {
  const lim = 0.03 * Math.min(1, (dIt || 100) / 100);
  const out = { ratios: spoolRatios(type, pw, etam), Tt4: NaN, Tt7: NaN };
  for (const st of controlStations(ctl)) {
    const sec = secOf(st), id = st.id, target = st.target;
    const T0 = comps[id * 4], K = comps[id * 4 + 1] || 1;
    if (sec.mdot > 0 && sec.Tt > T0 + 20 && target > T0 + 20) comps[id * 4 + 1] = clamp(K * clamp(Math.pow((target - T0) / (sec.Tt - T0), 0.3), 1 - lim, 1 + lim), 0.5, 2.5);
    out[id === 4 ? 'Tt4' : 'Tt7'] = sec.Tt;
  }
  return out;
}

/* Column and row mask of cfdSection at x over stream si (null: all fluid),
 * for the device monitor. */
function sectionRows(grid, zones, x, si) {
  const { W, H } = grid, i = colAt(grid, x), rows = new Uint8Array(H);
  for (let j = 0; j < H; j++) {
    const k = j * W + i;
    rows[j] = !zones.solid[k] && (si === null || zones.stream[k] === si) ? 1 : 0;
  }
  return { col: i, rows };
}

/* {mdot, Tt} of a section from the device monitor's sums (cfdSection's
 * division by the positive mass flow, 1 when there is none). */
const sectionFromSums = s => ({ mdot: s.mdot, Tt: s.wTt / (s.w || 1) });

/* Display fields of an axisymmetric state (cfdDerived, src/cfd_post.js:10-40):
 * region (0 exterior, 2 + stream), material, T, Tt, P, Pt, |V|, M, rho, Vx,
 * Vth (0), Vr, Y, entropy change ds against the freestream, and the azimuthal
 * vorticity du/dr - dv/dx by central differences over fluid neighbours.
 * disp: cut cells under half full show as solid (the mask is cached on zones
 * as dsolid). G: a previous result of the same size to refill, or null. */
const DERIVED_KEYS = ['T', 'Tt', 'P', 'Pt', 'V', 'M', 'rho', 'Vx', 'Vth', 'Vr', 'Y', 'vort', 'ds'];
function cfdDerived(grid, zones, Q, fl, G, disp) {
  const { W, H, N, rc } = grid;
  if (!G || G.N !== N) {
    G = { W, H, N, reg: new Uint8Array(N), mat: new Uint8Array(N) };
    for (const k of DERIVED_KEYS) G[k] = new Float32Array(N);
  }
  const S = disp ? (zones.dsolid || (zones.dsolid = Uint8Array.from(zones.solid, (v, k) => v || zones.phi[k] < 0.5 ? 1 : 0))) : zones.solid;
  const T0 = fl.T0, P0 = fl.P0;
  for (let k = 0; k < N; k++) {
    if (S[k]) {
      G.reg[k] = 1; G.mat[k] = zones.mat[k];
      for (const q of ['T', 'Tt', 'P', 'Pt', 'V', 'M', 'rho', 'Vx', 'Vr', 'Y', 'vort', 'ds']) G[q][k] = NaN;
      continue;
    }
    const st = zones.stream[k];
    G.reg[k] = st < 0 ? 0 : 2 + st;
    const r = Math.max(Q[0][k], 1e-4), u = Q[1][k] / r, v = Q[2][k] / r, Y = clamp(Q[4][k] / r, 0, 1), g = 1.4 - 0.07 * Y;
    const p = Math.max(1, (g - 1) * (Q[3][k] - 0.5 * r * (u * u + v * v)));
    const T = p / (r * RGAS), cp = g * RGAS / (g - 1), V2 = u * u + v * v, a = Math.sqrt(g * p / r);
    const Tt = T + V2 / (2 * cp);
    G.T[k] = T; G.P[k] = p; G.rho[k] = r; G.Vx[k] = u; G.Vr[k] = v; G.Vth[k] = 0;
    G.V[k] = Math.sqrt(V2); G.M[k] = Math.sqrt(V2) / a;
    G.Tt[k] = Tt; G.Pt[k] = p * Math.pow(Tt / T, g / (g - 1)); G.Y[k] = Y;
    G.ds[k] = cp * Math.log(T / T0) - RGAS * Math.log(p / P0);
  }
  for (let j = 0; j < H; j++) for (let i = 0; i < W; i++) {
    const k = j * W + i;
    if (S[k]) continue;
    const kl = i > 0 && !S[k - 1] ? k - 1 : k, kr = i < W - 1 && !S[k + 1] ? k + 1 : k;
    const kd = j > 0 && !S[k - W] ? k - W : k, ku = j < H - 1 && !S[k + W] ? k + W : k;
    const dxx = grid.xc[kr % W] - grid.xc[kl % W], drr = rc[(ku / W) | 0] - rc[(kd / W) | 0];
    const dvdx = dxx > 0 ? (G.Vr[kr] - G.Vr[kl]) / dxx : 0, dudr = drr > 0 ? (G.Vx[ku] - G.Vx[kd]) / drr : 0;
    G.vort[k] = dudr - dvdx;
  }
  return G;
}

/* Station averages of an engine flowpath (tools/stations.js:8-16): at each
 * station column, mass flow and mass-weighted Tt, Pt over the station's
 * stream, with the cell mass flow taken from the mean axial face mass flux.
 * list: an engine type ('tj', 'hbtf', 'lbtf') or [[id, x, stream], ...]. */
const STATIONS = {
  hbtf: [['2', 0.58, 'inlet'], ['13', 0.95, 'bypass'], ['25', 1.47, 'core'], ['3', 2.05, 'core'], ['4', 2.43, 'core'], ['45', 2.62, 'core'], ['5', 3.08, 'core'], ['19', 2.19, 'bypass']],
  tj: [['2', 0.36, 'core'], ['3', 1.55, 'core'], ['4', 2.21, 'core'], ['5', 2.62, 'core'], ['7', 4.29, 'core'], ['9', 4.89, 'core']],
  lbtf: [['2', 0.3, 'inlet'], ['13', 0.9, 'bypass'], ['3', 1.55, 'core'], ['4', 2.0, 'core'], ['5', 2.55, 'core'], ['6A', 3.0, 'ab'], ['9', 4.94, 'ab']],
};
const TJ_STATIONS = STATIONS.tj;
function stations(grid, zones, Q, mfx, order, list) {
  const { W, H, rc, dr } = grid, out = [];
  const L = typeof list === 'string' ? STATIONS[list] : (list || TJ_STATIONS);
  for (const [id, x, sn] of L) {
    const si = order.indexOf(sn), i = colAt(grid, x);
    let m = 0, tt = 0, pt = 0, wsum = 0;
    for (let j = 0; j < H; j++) {
      const k = j * W + i;
      if (zones.solid[k] || zones.stream[k] !== si) continue;
      const f = mfx[k] * 2 * Math.PI * rc[j] * dr[j];
      m += f;
      const q = cfdPrim(Q, k), cp = q.g * 287.05 / (q.g - 1), T = q.p / (q.r * 287.05), Tt = T + 0.5 * (q.u * q.u + q.v * q.v) / cp, Pt = q.p * Math.pow(Tt / T, q.g / (q.g - 1));
      const wk = Math.max(f, 0);
      tt += wk * Tt; pt += wk * Pt; wsum += wk;
    }
    out.push({ id, x, mdot: m, Tt: tt / wsum, Pt: pt / wsum });
  }
  return out;
}

module.exports = { RGAS, clamp, cfdPrim, cfdDerived, DERIVED_KEYS, cfdForces, colAt, cfdSection, spoolRatios, cfdControl, controlStations,
  controlFromSections, sectionRows, sectionFromSums, stations, STATIONS, TJ_STATIONS };
