'use strict';
/* ============================================================================
 * field_model.js — the engine field model on the device, for the two places
 * the solvers read it:
 *   initialState(grid, zones, fm, uniform)   the axisymmetric solver's start
 *                                            (src/cfd_grid.js:279-305)
 *   fixedStates(fm, fs, M, alphaDeg)         the 3D solver's freestream and its
 *                                            inlet-sink / jet fixed states
 *                                            (CFD3.applyCond, src/air_cfd.js:521-538)
 * FieldHIP is the FieldModel object (src/field.js:255-488) on the device:
 * built from a field model description — the constructor's products (stream
 * knots with their Mach numbers, processed exits, inlet face velocity, hot
 * gas) and the engine geometry it samples, as its owner hands them over —
 * it samples points (FieldModel.prototype.sample) and fills initial states,
 * one device thread per point / cell (axi_field_*, axi_initial_state).
 * ========================================================================== */
const { loadAddon } = require('./hip_solver.js');

const STREAM_NAME = { inlet: 0, bypass: 1, core: 2, ab: 3 };
const ENGINE_TYPE = { tj: 0, hbtf: 1, lbtf: 2 };
const FM_HEAD = 240;            // AXI_FM_HEAD (include/axi.h)
const SAMPLE_KEYS = ['T', 'P', 'Tt', 'Pt', 'V', 'Vx', 'Vth', 'M', 'a', 'rho', 'g', 'cp', 'R'];

/* The flat table of axi_field_create (layout AXI_FM_* in include/axi.h). */
function fieldTable(d) {
  const T = new Array(FM_HEAD).fill(0);
  const num = v => (v === null || v === undefined ? NaN : +v);
  const put = vals => { const off = T.length; for (const v of vals) T.push(+v); return off; };
  const poly = pts => { const off = T.length; for (const p of pts) T.push(+p[0], +p[1]); return [off, pts.length]; };
  const setAt = (i, vals) => { for (let q = 0; q < vals.length; q++) T[i + q] = +vals[q]; };
  const order = d.order;
  setAt(0, [1, ENGINE_TYPE[d.type], d.s, order.length]);
  order.forEach((k, i) => {
    const st = d.streams[k], lo = poly(st.lo), hi = poly(st.hi);
    const kn = put([].concat(...d.kn[k]));
    setAt(4 + 10 * i, [STREAM_NAME[k], st.x0, st.x1, lo[0], lo[1], hi[0], hi[1], kn, d.kn[k].length, 0]);
  });
  T[44] = d.bodies.length;
  d.bodies.forEach((b, i) => {
    let x0 = Infinity, x1 = -Infinity, r0 = Infinity, r1 = -Infinity;
    for (const p of b.poly) { x0 = Math.min(x0, p[0]); x1 = Math.max(x1, p[0]); r0 = Math.min(r0, p[1]); r1 = Math.max(r1, p[1]); }
    const pl = poly(b.poly);
    setAt(45 + 6 * i, [x0, x1, r0, r1, pl[0], pl[1]]);
  });
  const cb = d.comb, li = poly(cb.linerIn), lo = poly(cb.linerOut);
  setAt(141, [order.indexOf(cb.stream), cb.x0, cb.xd, cb.x1, cb.turbEnd, cb.n, cb.cannular ? 1 : 0, li[0], li[1], lo[0], lo[1]]);
  setAt(152, d.mixer ? [1, d.mixer.x, d.mixer.coreFrac, d.mixer.lobes] : [0, 0, 0, 0]);
  if (d.ab) { const rg = put(d.ab.rings); setAt(156, [1, order.indexOf(d.ab.stream), d.ab.xFH, d.ab.x1, rg, d.ab.rings.length]); }
  const c = d.c;
  setAt(162, [num(c.Tt3), num(c.Tt4), num(c.Tt5), num(c.Tt6A), num(c.Tt7), num(c.Tt13), c.abOn ? 1 : 0]);
  setAt(169, [d.fl.V0, d.fl.T0, d.fl.P0, d.fl.Tt0, d.fl.Pt0]);
  const ol = poly(d.outer);
  setAt(174, [d.lip.x, d.lip.r, d.lip.rOut, ol[0], ol[1], d.outerEnd]);
  setAt(180, [d.cold.g, d.cold.cp]);
  T[182] = d.exits.length;
  d.exits.forEach((e, q) => setAt(183 + 12 * q, [e.x, e.R, e.spread, e.Qm, e.Qh, e.under ? 1 : 0, e.Ls, e.P, e.Pt, e.g, e.cp, e.dTt]));
  const hg = d.hotGas;
  setAt(231, [d.Uface, d.maxdT, hg ? 1 : 0, hg ? hg.g : 0, hg ? hg.cp : 0]);
  return Float64Array.from(T);
}

class FieldHIP {
  constructor(desc) {
    this.desc = desc;
    this.order = desc.order.slice();
    this.table = fieldTable(desc);
    this.h = loadAddon().fieldCreate(this.table);        // throws when the table or the device is unusable
  }
  /* n points: Float64Array(14 n) of region (0 outside, 1 solid, 2 + stream) and SAMPLE_KEYS */
  sampleMany(x, r, th) {
    const n = x.length, out = new Float64Array(14 * n);
    loadAddon().fieldSample(this.h, Float64Array.from(x), Float64Array.from(r), th ? Float64Array.from(th) : new Float64Array(n), out);
    return out;
  }
  /* FieldModel.prototype.sample(x, r, th, out) for one point: the same fields */
  sample(x, r, th, out) {
    const v = this.sampleMany([x], [r], [th || 0]);
    out = out || {};
    const reg = v[0];
    out.x = x; out.r = r; out.th = th || 0;
    out.solid = reg === 1;
    out.region = reg === 1 ? 'solid' : reg === 0 ? 'ext' : 'stream';
    out.stream = reg >= 2 ? this.order[reg - 2] : null;
    if (!out.solid) SAMPLE_KEYS.forEach((k, q) => { out[k] = v[1 + q]; });
    return out;
  }
  dispose() { if (this.h) { loadAddon().fieldDispose(this.h); this.h = null; } }
}

/* initialState(grid, zones, spec, model, fl, uniform) (src/cfd_grid.js:279-305)
 * with the model on the device: Q as five Float64Array(N), [rho, rho u, 0, E, rho Y]. */
function initialState(grid, zones, fm, uniform) {
  const N = grid.W * grid.H, Q = new Float64Array(5 * N);
  const stream = zones.stream instanceof Int8Array ? zones.stream : Int8Array.from(zones.stream);
  const solid = zones.solid instanceof Uint8Array ? zones.solid : Uint8Array.from(zones.solid);
  loadAddon().initialState(fm.h, Float64Array.from(grid.xf), Float64Array.from(grid.rf), solid, stream, !!uniform, Q);
  return [0, 1, 2, 3, 4].map(q => Q.subarray(q * N, (q + 1) * N));
}

/* wall radius of a polyline at x: flat beyond its ends, linear between the
 * two points that bracket x (the reference's Line.at, src/geometry.js:27-34) */
function wallAt(pts, x) {
  const n = pts.length;
  if (x <= pts[0][0]) return pts[0][1];
  if (x >= pts[n - 1][0]) return pts[n - 1][1];
  let a = 0, b = n - 1;
  while (b - a > 1) { const m = (a + b) >> 1; if (pts[m][0] <= x) a = m; else b = m; }
  const dx = pts[b][0] - pts[a][0];
  return dx < 1e-12 ? pts[b][1] : pts[a][1] + (pts[b][1] - pts[a][1]) * (x - pts[a][0]) / dx;
}

/* CFD3.applyCond (src/air_cfd.js:521-538): the freestream F of the Mach
 * number and angle of attack, and the three fixed states [sink, bypass jet,
 * core jet] from the engine's field model sampled just behind the fan face
 * (station 2) and just ahead of each nozzle exit, in the solver's units
 * (density and pressure over the flight state's, velocity over its sound
 * speed, gamma 1.4); a point that falls in a solid, or an absent exit, takes
 * F. fs: the flight state {P, T, a}. fm: a FieldHIP. */
function fixedStates(fm, fs, M, alphaDeg) {
  const CG = 1.4, a = (alphaDeg || 0) * Math.PI / 180, u = M * Math.cos(a), w = M * Math.sin(a);
  const F = [1, u, 0, w, 1 / (CG * (CG - 1)) + 0.5 * M * M];
  const d = fm.desc, s = d.s;
  const pts = [], who = [];
  const st2 = (d.stations || []).find(q => q.id === '2');
  const x2 = st2 ? st2.x : d.lip.x + 0.1 * s;
  const rin = d.streams.inlet || d.streams.core;
  pts.push([x2 + 0.01 * s, rin ? 0.5 * (wallAt(rin.lo, x2) + wallAt(rin.hi, x2)) : 0.3 * d.lip.r]); who.push('sink');
  const exB = d.exits.find(q => q.key === 'bypass'), exC = d.exits.find(q => q.key !== 'bypass');
  if (exB) { pts.push([exB.x - 0.02 * s, 0.5 * (exB.R + exB.Rin)]); who.push('jB'); }
  if (exC) { pts.push([exC.x - 0.02 * s, 0.5 * (exC.R + (exC.Rin || 0))]); who.push('jC'); }
  const v = fm.sampleMany(pts.map(p => p[0]), pts.map(p => p[1]));
  const fix = { sink: F, jB: F, jC: F };
  who.forEach((k, p) => {
    if (v[14 * p] === 1) return;                    // inside a solid: the freestream stands in
    const P = v[14 * p + 2], T = v[14 * p + 1], Vx = v[14 * p + 6];
    const rr = (P / fs.P) / (T / fs.T), pp = (P / fs.P) / CG, uu = Vx / fs.a;
    fix[k] = [rr, rr * uu, 0, 0, pp / (CG - 1) + 0.5 * rr * uu * uu];
  });
  return { F, fix: [fix.sink, fix.jB, fix.jC] };
}

module.exports = { FieldHIP, fieldTable, initialState, fixedStates, wallAt, SAMPLE_KEYS };
