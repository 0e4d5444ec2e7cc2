'use strict';
/* ============================================================================
 * cfd_grid.js — setup of the axisymmetric engine solver: the banded (x, r)
 * grid, the cut-cell / stream / blade-row / heat-source zones and the
 * component constants (buildCFDGrid, buildZones, updateHeatZones,
 * streamMassFlows, compParams; src/cfd_grid.js:11-276), producing the typed
 * arrays AxiHIP takes. The state transfer between grids of the grid
 * sequencing (interpState, src/app_cfd.js:321-358) runs on the device
 * (AxiHIP.transferFrom).
 *
 * Input: an engine description as data (`axispec`: stream walls and liners as
 * polyline points, body polygons, blade rows, combustor / afterburner extents,
 * nozzle exits, the 1D cycle's temperatures and mass flows, component
 * efficiencies), i.e. what the reference's buildSpec (src/geometry.js) and
 * solveEngine (src/physics.js) hand to its setup. The initial state
 * (initialState, src/cfd_grid.js:279-305) samples the reference's 1D field
 * model and stays an input.
 *
 * Every expression keeps the reference's operation order: the zones are
 * compared byte for byte with the reference's (Float32 arrays of doubles).
 * ========================================================================== */

const GRID_LEVELS = { xcoarse: { nx: 120, nr: 70 }, coarse: { nx: 190, nr: 110 }, medium: { nx: 320, nr: 200 }, fine: { nx: 460, nr: 290 } };   // src/cfd_grid.js:8
const COMP_ID = { fan: 1, lpc: 2, hpc: 3, comp: 3, comb: 4, hpt: 5, turb: 5, lpt: 6, ab: 7 };                                                      // src/cfd_grid.js:6
const MAT_ID = { nacelle: 1, casing: 1, hub: 2, cowl: 3, liner: 4, fh: 4 };                                                                        // src/cfd_grid.js:87
const DEG = Math.PI / 180;

const clamp = (v, a, b) => (v < a ? a : v > b ? b : v);
const lerp = (a, b, t) => a + (b - a) * t;
function smoothstep(e0, e1, x) { const t = clamp((x - e0) / (e1 - e0), 0, 1); return t * t * (3 - 2 * t); }

/* Piecewise-linear wall r(x) through points with non-decreasing x, constant
 * beyond the ends (Line.at, src/geometry.js:27-34). */
function wall(pts) {
  const n = pts.length, xs = pts.map(p => p[0]), rs = pts.map(p => p[1]);
  return {
    at(x) {
      if (x <= xs[0]) return rs[0];
      if (x >= xs[n - 1]) return rs[n - 1];
      let a = 0, b = n - 1;
      while (b - a > 1) { const m = (a + b) >> 1; if (xs[m] <= x) a = m; else b = m; }
      const w = xs[b] - xs[a];
      return w < 1e-12 ? rs[b] : rs[a] + (rs[b] - rs[a]) * (x - xs[a]) / w;
    },
  };
}

function bbox(poly) {
  let x0 = Infinity, x1 = -Infinity, r0 = Infinity, r1 = -Infinity;
  for (const [x, r] of poly) { x0 = Math.min(x0, x); x1 = Math.max(x1, x); r0 = Math.min(r0, r); r1 = Math.max(r1, r); }
  return [x0, x1, r0, r1];
}
/* even-odd crossing test (src/geometry.js:61-68) */
function insidePoly(x, r, poly) {
  let inside = false;
  for (let i = 0, j = poly.length - 1; i < poly.length; j = i++) {
    const xi = poly[i][0], ri = poly[i][1], xj = poly[j][0], rj = poly[j][1];
    if ((ri > r) !== (rj > r) && x < (xj - xi) * (r - ri) / (rj - ri) + xi) inside = !inside;
  }
  return inside;
}

/* Polygon edges [xi, ri, xj, rj] (in insidePoly's order, so the crossing
 * abscissa is computed from the same operands) in NB bins of r over the
 * bounding box; an edge sits in every bin its r-range touches. An edge whose
 * ends straddle r has min(ri, rj) <= r < max(ri, rj), so it is in r's bin and
 * the even-odd count over a bin equals the count over the whole polygon. */
const NB = 64;
function binEdges(poly, bb) {
  const r0 = bb[2], sc = NB / Math.max(1e-30, bb[3] - bb[2]);
  const bin = v => Math.min(NB - 1, Math.max(0, Math.floor((v - r0) * sc)));
  const bins = [];
  for (let b = 0; b < NB; b++) bins.push([]);
  for (let i = 0, j = poly.length - 1; i < poly.length; j = i++) {
    const xi = poly[i][0], ri = poly[i][1], xj = poly[j][0], rj = poly[j][1];
    for (let b = bin(Math.min(ri, rj)); b <= bin(Math.max(ri, rj)); b++) bins[b].push(xi, ri, xj, rj);
  }
  return { r0, sc, e: bins.map(a => Float64Array.from(a)) };
}
/* insidePoly through the binned edges: the same tests on the edges that can toggle */
function insideBinned(x, r, B) {
  const e = B.e[Math.min(NB - 1, Math.max(0, Math.floor((r - B.r0) * B.sc)))];
  let inside = false;
  for (let t = 0; t < e.length; t += 4) {
    const xi = e[t], ri = e[t + 1], xj = e[t + 2], rj = e[t + 3];
    if ((ri > r) !== (rj > r) && x < (xj - xi) * (r - ri) / (rj - ri) + xi) inside = !inside;
  }
  return inside;
}

/* The engine description with its walls as functions and body bounding boxes. */
function engine(spec) {
  const streams = {};
  for (const k of spec.order) { const st = spec.streams[k]; streams[k] = { x0: st.x0, x1: st.x1, lo: wall(st.lo), hi: wall(st.hi) }; }
  const bodies = spec.bodies.map(b => { const bb = bbox(b.poly); return { key: b.key, mat: b.mat, poly: b.poly, bb, bins: binEdges(b.poly, bb) }; });
  const comb = Object.assign({}, spec.comb, { linerIn: wall(spec.comb.linerIn), linerOut: wall(spec.comb.linerOut) });
  return Object.assign({}, spec, { streams, bodies, comb });
}

/* Region of point (x, r): the first stream whose walls enclose it, else a
 * solid body, else the exterior (FieldModel.classify, src/field.js:296-309). */
function classify(E, x, r) {
  for (const key of E.order) {
    const st = E.streams[key];
    if (x < st.x0 || x > st.x1) continue;
    const lo = st.lo.at(x), hi = st.hi.at(x);
    if (r >= lo && r <= hi) return { kind: 'stream', key };
  }
  for (const b of E.bodies) {
    const bb = b.bb;
    if (x < bb[0] || x > bb[1] || r < bb[2] || r > bb[3]) continue;
    if (insidePoly(x, r, b.poly)) return { kind: 'solid', mat: b.mat };
  }
  return { kind: 'ext' };
}
/* classify on one grid column x, as a code: stream index >= 0, -1 exterior,
 * -2 - b for body b. Wall radii and the bodies whose x-range holds x are taken
 * once per column; the tests and their order are classify's. */
function columnClassifier(E, x) {
  const lo = [], hi = [], id = [];
  E.order.forEach((key, s) => {
    const st = E.streams[key];
    if (x < st.x0 || x > st.x1) return;
    lo.push(st.lo.at(x)); hi.push(st.hi.at(x)); id.push(s);
  });
  const bs = [], bi = [];
  E.bodies.forEach((b, n) => { if (!(x < b.bb[0] || x > b.bb[1])) { bs.push(b); bi.push(n); } });
  return r => {
    for (let s = 0; s < id.length; s++) if (r >= lo[s] && r <= hi[s]) return id[s];
    for (let n = 0; n < bs.length; n++) {
      const bb = bs[n].bb;
      if (r < bb[2] || r > bb[3]) continue;
      if (insideBinned(x, r, bs[n].bins)) return -2 - bi[n];
    }
    return -1;
  };
}

/* Faces of one axis: refinement bands [a, b, h], growth g outside them, cap
 * hmax; a last cell thinner than 0.35 h is merged (src/cfd_grid.js:11-26). */
function axisFaces(lo, hi, bands, g, hmax) {
  const spacing = x => {
    let h = hmax;
    for (const [a, b, hb] of bands) { const d = x < a ? a - x : x > b ? x - b : 0; h = Math.min(h, hb + (g - 1) * d); }
    return h;
  };
  const f = [lo];
  let x = lo;
  while (x < hi - 1e-9) {
    const h = spacing(x + 0.5 * spacing(x));
    x += h;
    if (hi - x < 0.35 * h) x = hi;
    f.push(Math.min(x, hi));
  }
  return Float64Array.from(f);
}

/* Grid of a level (src/cfd_grid.js:28-43). */
function buildCFDGrid(spec, level) {
  const L = GRID_LEVELS[level] || GRID_LEVELS.medium;
  const s = spec.s, len = spec.view.len, Rm = spec.view.Rmax;
  const xb = spec.domain.xb, rb = spec.domain.rb;
  const hx = len / L.nx, hr = Rm / L.nr;
  const xf = axisFaces(xb[0], xb[xb.length - 1], [[-0.04 * len, 1.04 * len, hx]], 1.045, 16 * hx);
  // the high-bypass turbofan refines the (narrow) core duct's radius band
  const bands = spec.type === 'hbtf' ? [[0, 1.06 * Rm, hr * 1.5], [0.11 * s, 0.43 * s, hr * 0.55]] : [[0, 1.06 * Rm, hr]];
  const rf = axisFaces(0, rb[rb.length - 1], bands, 1.05, 14 * hr);
  const W = xf.length - 1, H = rf.length - 1;
  const xc = new Float64Array(W), dx = new Float64Array(W), rc = new Float64Array(H), dr = new Float64Array(H);
  for (let i = 0; i < W; i++) { xc[i] = 0.5 * (xf[i] + xf[i + 1]); dx[i] = xf[i + 1] - xf[i]; }
  for (let j = 0; j < H; j++) { rc[j] = 0.5 * (rf[j] + rf[j + 1]); dr[j] = rf[j + 1] - rf[j]; }
  return { W, H, N: W * H, xf, rf, xc, rc, dx, dr, level };
}

/* Solid radius intervals along the vertical line x: inside a body polygon and
 * outside every stream (src/cfd_grid.js:46-72). */
function solidIntervals(E, bodies, x) {
  let segs = [];
  for (const b of bodies) {
    const bb = b.bb;
    if (x < bb[0] || x > bb[1]) continue;
    const P = b.poly, cr = [];
    for (let i = 0, j = P.length - 1; i < P.length; j = i++) {
      const x1 = P[j][0], r1 = P[j][1], x2 = P[i][0], r2 = P[i][1];
      if ((x1 > x) !== (x2 > x)) cr.push(r1 + (r2 - r1) * (x - x1) / (x2 - x1));
    }
    cr.sort((p, q) => p - q);
    for (let k = 0; k + 1 < cr.length; k += 2) segs.push([cr[k], cr[k + 1]]);
  }
  for (const key of E.order) {
    const st = E.streams[key];
    if (x < st.x0 || x > st.x1) continue;
    const lo = st.lo.at(x), hi = st.hi.at(x), kept = [];
    for (const [a, c] of segs) {
      if (c <= lo || a >= hi) { kept.push([a, c]); continue; }
      if (a < lo) kept.push([a, lo]);
      if (c > hi) kept.push([hi, c]);
    }
    segs = kept;
  }
  return segs;
}
/* solid length of [r0, r1] (r-weighted: the area 0.5 (hi^2 - lo^2)) */
function solidLen(iv, r0, r1, rw) {
  let s = 0;
  for (const [a, c] of iv) { const lo = Math.max(a, r0), hi = Math.min(c, r1); if (hi > lo) s += rw ? 0.5 * (hi * hi - lo * lo) : hi - lo; }
  return s;
}
function solidAt(iv, r) { for (const [a, c] of iv) if (r > a && r < c) return true; return false; }

/* Cycle mass flow of each stream (src/cfd_grid.js:236-240). */
function streamMassFlows(E, op) {
  if (E.type === 'tj') return { core: op.mdot, core4: op.mdot * (1 + op.f) };
  const m4 = op.mcore * (1 + op.f);
  return { inlet: op.mdot, core: op.mcore, core4: m4, bypass: op.mbyp, ab: op.m7 || m4 + op.mbyp };
}

/* Heat-source targets (combustor exit-temperature profile, afterburner flame
 * front) and blade-row design mass fluxes (src/cfd_grid.js:194-234). */
function updateHeatZones(grid, zones, E, op) {
  const { W, H, xc, rc } = grid, s = E.s;
  const { comp, tgt, rate, yrate, solid } = zones;
  const cb = E.comb, Tt3 = op.Tt3, Tt4 = op.Tt4, dT = Math.max(1, Tt4 - Tt3);
  const Tpz = Math.min(Tt3 + 1.9 * dT, 2480, Tt3 + 1800);
  const tauC = 0.1 * (cb.x1 - cb.xd);
  const ab = E.ab, tauA = ab ? 0.08 * (ab.x1 - ab.xFH) : 1;
  const mS = streamMassFlows(E, op);
  for (let j = 0; j < H; j++) for (let i = 0; i < W; i++) {
    const k = j * W + i;
    tgt[k] = 0; rate[k] = 0; yrate[k] = 0; zones.mfd[k] = 0;
    if (solid[k]) continue;
    const x = xc[i], r = rc[j];
    if (comp[k] && comp[k] !== 4 && comp[k] !== 7) {
      const key = E.order[zones.stream[k]], st = E.streams[key];
      const lo = st.lo.at(x), hi = st.hi.at(x), A = Math.PI * (hi * hi - lo * lo);
      const m = key === 'core' && comp[k] >= 5 ? mS.core4 : mS[key];
      zones.mfd[k] = m / Math.max(A, 1e-6);
    }
    if (comp[k] === 4) {
      const st = E.streams[cb.stream], lo = st.lo.at(x), hi = st.hi.at(x);
      const sN = (x - cb.xd) / (cb.x1 - cb.xd), psi = clamp((r - lo) / Math.max(1e-6, hi - lo), 0, 1);
      const li = cb.linerIn.at(x), lu = cb.linerOut.at(x);
      const pl = (r - li) / Math.max(1e-6, lu - li);
      const Tline = sN < 0.3 ? Tt3 + (Tpz - Tt3) * smoothstep(0.0, 0.26, sN) : Tt4 + (Tpz - Tt4) * Math.exp(-3.4 * (sN - 0.3) / 0.7);
      const gP = pl < 0 || pl > 1 ? 0 : 0.25 + 0.75 * Math.exp(-(((pl - 0.5) / 0.26) ** 2));
      const gE = 1 + 0.3 * (Math.sin(Math.PI * psi) - 2 / Math.PI);
      const G = lerp(gP, gE, smoothstep(0.3, 1, sN));
      tgt[k] = Tt3 + (Tline - Tt3) * G;
      rate[k] = 1 / tauC;
      yrate[k] = G > 0.35 ? 1 / tauC : 0;
    } else if (comp[k] === 7 && op.abOn) {
      const t = (x - ab.xFH) / (ab.x1 - ab.xFH);
      const spread = (x - ab.xFH) * Math.tan(7 * DEG) + 0.006 * s;
      let F = 0;
      for (const rg of ab.rings) F = Math.max(F, 1 - smoothstep(spread * 0.5, spread, Math.abs(r - rg)));
      F = Math.max(F, smoothstep(0.5, 1.0, t));
      tgt[k] = op.Tt7; rate[k] = F / tauA; yrate[k] = F / tauA;
    }
  }
}

/* The cut-cell pass of buildZones over columns i0..i1-1 (src/cfd_grid.js:81-128):
 * volume fraction, east / north open fractions, split cells, stream and
 * material per cell, into out = {phi, ae, an, split, splitFr, stream, mat}.
 * Columns are independent, so a range can run on any thread. One column per
 * call: the engine optimises the column function from its own type feedback. */
function zoneColumns(E, grid, out, i0, i1) {
  const { W, H, xc, rc, dx, xf, rf } = grid;
  const { phi, ae, an, split, splitFr, stream, mat } = out;
  const bodies = E.bodies.filter(b => b.mat !== 'liner' && b.mat !== 'fh');
  const NS = 8;
  const matOf = E.bodies.map(b => MAT_ID[b.mat] || 1);
  const column = i => {
    const ivF = solidIntervals(E, bodies, xf[i + 1]);
    const ivs = [];
    for (let q = 0; q < NS; q++) ivs.push(solidIntervals(E, bodies, xf[i] + dx[i] * (q + 0.5) / NS));
    const ivC = solidIntervals(E, bodies, xc[i]);
    const cls = columnClassifier(E, xc[i]);
    for (let j = 0; j < H; j++) {
      const k = j * W + i, r0 = rf[j], r1 = rf[j + 1], vr = 0.5 * (r1 * r1 - r0 * r0);
      let f = 0;
      for (const iv of ivs) f += 1 - solidLen(iv, r0, r1, true) / vr;
      phi[k] = f / NS;
      ae[k] = 1 - solidLen(ivF, r0, r1, false) / (r1 - r0);
      let open = 0;
      for (const iv of ivs) if (!solidAt(iv, r1)) open++;
      an[k] = open / NS;
      // stream of the cell: from a fluid point of the centre line
      let rs = rc[j];
      if (solidAt(ivC, rs)) { for (let q = 1; q <= 6; q++) { const rr = r0 + (r1 - r0) * q / 7; if (!solidAt(ivC, rr)) { rs = rr; break; } } }
      // a thin wall inside the cell (touching neither face) on at least half the
      // sub-lines: keep the larger side and close the other one's radial face
      let nIn = 0, aS = 0, cS = 0;
      for (const iv of ivs) for (const [a, c] of iv) if (a > r0 + 1e-9 && c < r1 - 1e-9) { nIn++; aS += a; cS += c; break; }
      if (nIn * 2 >= NS) {
        const a = aS / nIn, c = cS / nIn, lo = 0.5 * (a * a - r0 * r0) / vr, up = 0.5 * (r1 * r1 - c * c) / vr;
        const keepLow = lo >= up;
        split[k] = keepLow ? 1 : 2;
        splitFr[k] = keepLow ? (a - r0) / (r1 - r0) : (r1 - c) / (r1 - r0);
        phi[k] = keepLow ? lo : up;
        ae[k] = Math.min(ae[k], splitFr[k]);
        rs = keepLow ? 0.5 * (r0 + a) : 0.5 * (c + r1);
      }
      const cc = cls(rc[j]), cl = rs === rc[j] ? cc : cls(rs);
      stream[k] = cl >= 0 ? cl : -1;
      mat[k] = cc <= -2 ? matOf[-2 - cc] : 1;
    }
  };
  for (let i = i0; i < i1; i++) column(i);
}

/* Zones (src/cfd_grid.js:81-191): cut cells (volume fraction phi, east / north
 * face open fractions ae / an from 8 axial sub-lines of solid intervals),
 * thin-wall split cells, small cut cells merged into solids, stream index per
 * cell (with a breadth-first fill of fluid cells no stream contains), blade-row
 * weights, combustor / afterburner cells, heat targets, mixing length.
 * `pre`: the column pass's arrays for every column, computed elsewhere (the
 * worker-thread build); they are completed in place. */
function buildZones(grid, spec, op, pre) {
  const E = engine(spec);
  const { W, H, N, xc, dx, xf } = grid, s = E.s;
  const out = pre || { phi: new Float32Array(N), ae: new Float32Array(N), an: new Float32Array(N), split: new Uint8Array(N),
    splitFr: new Float32Array(N), stream: new Int8Array(N), mat: new Uint8Array(N) };
  if (!pre) zoneColumns(E, grid, out, 0, W);
  const { phi, ae, an, split, splitFr, stream, mat } = out;
  const solid = new Uint8Array(N), comp = new Uint8Array(N), wgt = new Float32Array(N), lmix = new Float32Array(N);
  // each pass below is a function of its own (optimised from its own type feedback)
  const splitFaces = () => {
    for (let k = 0; k < N; k++) if (split[k]) {
      const i = k % W, j = (k / W) | 0;
      if (split[k] === 1) an[k] = 0; else if (j > 0) an[k - W] = 0;
      if (i > 0) ae[k - 1] = Math.min(ae[k - 1], splitFr[k]);
    }
  };
  splitFaces();
  const merge = () => { for (let k = 0; k < N; k++) if (phi[k] < 0.08) { phi[k] = 0; solid[k] = 1; stream[k] = -2; } };
  merge();
  const wallFaces = () => {
    for (let j = 0; j < H; j++) for (let i = 0; i < W; i++) {
      const k = j * W + i;
      if (solid[k] || i === W - 1 || solid[k + 1]) ae[k] = i === W - 1 && !solid[k] ? 1 : 0;
      if (solid[k] || j === H - 1 || solid[k + W]) an[k] = j === H - 1 && !solid[k] ? 1 : 0;
      ae[k] = Math.min(ae[k], 1); an[k] = Math.min(an[k], 1);
    }
  };
  wallFaces();
  // fluid cells in no stream: those connected to the outer boundary are
  // exterior; the rest take the stream of the nearest stream cell through open
  // faces (two multi-source breadth-first sweeps, directions E W N S)
  const fill = () => {
    const open = (k, d) => d === 0 ? ae[k] > 0 : d === 1 ? ae[k - 1] > 0 : d === 2 ? an[k] > 0 : an[k - W] > 0;
    const nb = (k, d) => {
      const i = k % W, j = (k / W) | 0;
      if (d === 0) return i < W - 1 ? k + 1 : -1;
      if (d === 1) return i > 0 ? k - 1 : -1;
      if (d === 2) return j < H - 1 ? k + W : -1;
      return j > 0 ? k - W : -1;
    };
    const ext = new Uint8Array(N);
    const sweep = (front, mark) => {
      while (front.length) {
        const next = [];
        for (const k of front) for (let d = 0; d < 4; d++) {
          const m = nb(k, d);
          if (m < 0 || solid[m] || ext[m] || stream[m] >= 0 || !open(k, d)) continue;
          mark(m, k); next.push(m);
        }
        front = next;
      }
    };
    let front = [];
    for (let k = 0; k < N; k++) {
      const i = k % W, j = (k / W) | 0;
      if (!solid[k] && stream[k] < 0 && (i === 0 || i === W - 1 || j === H - 1)) { ext[k] = 1; front.push(k); }
    }
    sweep(front, m => { ext[m] = 1; });
    front = [];
    for (let k = 0; k < N; k++) if (!solid[k] && stream[k] >= 0) front.push(k);
    sweep(front, (m, k) => { stream[m] = stream[k]; });
  };
  fill();
  // blade rows: rotors only do work; weight = axial overlap / (row length x rotors of the component)
  const rotors = E.rows.filter(r => r.rotor);
  const nRot = {};
  for (const r of rotors) nRot[r.comp] = (nRot[r.comp] || 0) + 1;
  const row = (r, id, si) => {
    for (let i = 0; i < W; i++) {
      const ov = Math.min(r.x1, xf[i + 1]) - Math.max(r.x0, xf[i]);
      if (ov <= 0) continue;
      const w = (ov / dx[i]) / ((r.x1 - r.x0) * nRot[r.comp]);
      for (let j = 0; j < H; j++) { const k = j * W + i; if (!solid[k] && stream[k] === si) { comp[k] = id; wgt[k] += w; } }
    }
  };
  for (const r of rotors) {
    const id = COMP_ID[r.comp];
    if (!id) continue;
    row(r, id, E.order.indexOf(r.stream));
  }
  // combustor and afterburner cells
  const span = (xa, xb, si, id) => {
    for (let i = 0; i < W; i++) {
      if (xc[i] < xa || xc[i] > xb) continue;
      for (let j = 0; j < H; j++) { const k = j * W + i; if (!solid[k] && stream[k] === si) comp[k] = id; }
    }
  };
  const cb = E.comb;
  span(cb.xd, cb.x1, E.order.indexOf(cb.stream), 4);
  if (E.ab) span(E.ab.xFH, E.ab.x1, E.order.indexOf(E.ab.stream), 7);
  const zones = { solid, mat, stream, comp, wgt, lmix, tgt: new Float32Array(N), rate: new Float32Array(N), yrate: new Float32Array(N),
    mfd: new Float32Array(N), phi, ae, an };
  updateHeatZones(grid, zones, E, op);
  // Prandtl mixing length: 0.025 / 0.07 of the duct height in streams, a jet
  // shear layer growing from the first exit outside
  const xe0 = Math.min(...E.exitsGeom.map(e => e.x)), Rj = Math.max(...E.exitsGeom.map(e => e.R));
  const mixRow = j => {
    for (let i = 0; i < W; i++) {
      const k = j * W + i;
      if (solid[k]) continue;
      const x = xc[i];
      if (stream[k] >= 0) {
        const st = E.streams[E.order[stream[k]]];
        const h = Math.max(1e-4, st.hi.at(x) - st.lo.at(x));
        lmix[k] = (comp[k] === 4 || comp[k] === 7 ? 0.07 : 0.025) * h;
      } else {
        const dxe = x - xe0;
        lmix[k] = dxe > 0 ? 0.075 * (0.05 * Rj + 0.12 * dxe) : 0.004 * s;
      }
    }
  };
  for (let j = 0; j < H; j++) mixRow(j);
  return zones;
}

/* Component constants: per id [ln tau (total-temperature ratio), polytropic
 * efficiency, 1 compression / 2 expansion, 0]; heat sources [base Tt, gain K]
 * (src/cfd_grid.js:260-276). */
function compParams(type, op, eff) {
  const c = new Float64Array(8 * 4);
  const set = (id, tau, ek, ty) => { c[id * 4] = Math.log(tau); c[id * 4 + 1] = ek; c[id * 4 + 2] = ty; };
  c[4 * 4] = op.Tt3; c[4 * 4 + 1] = 1; c[7 * 4] = op.Tt6 || op.Tt5 || 0; c[7 * 4 + 1] = 1;
  if (type === 'hbtf') {
    set(1, op.Tt13 / op.Tt2, eff.ef, 1); set(2, op.Tt25 / op.Tt13, eff.ec, 1); set(3, op.Tt3 / op.Tt25, eff.ec, 1);
    set(5, op.Tt45 / op.Tt4, eff.et, 2); set(6, op.Tt5 / op.Tt45, eff.et, 2);
  } else if (type === 'tj') {
    set(3, op.Tt3 / op.Tt2, eff.ec, 1); set(5, op.Tt5 / op.Tt4, eff.et, 2);
  } else {
    set(1, op.Tt13 / op.Tt2, eff.ef, 1); set(3, op.Tt3 / op.Tt13, eff.ec, 1);
    set(5, op.Tt45 / op.Tt4, eff.et, 2); set(6, op.Tt5 / op.Tt45, eff.et, 2);
  }
  return c;
}

/* Whole setup from an engine description: grid, zones, component constants. */
function axiSetup(spec, level) {
  const grid = buildCFDGrid(spec, level || spec.level);
  const zones = buildZones(grid, spec, spec.op);
  const comps = compParams(spec.type, spec.op, spec.eff);
  return { grid, zones, comps };
}

module.exports = { GRID_LEVELS, COMP_ID, axisFaces, buildCFDGrid, engine, zoneColumns, buildZones, updateHeatZones, streamMassFlows, compParams, axiSetup };
