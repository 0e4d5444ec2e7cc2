'use strict';
/* ============================================================================
 * cfd_axi.js — AxiHIP: the axisymmetric engine solver object of the
 * reference's protocol (GPUSolver / CPUSolver, src/cfd_gpu.js:207-323,
 * src/cfd_cpu.js:50-214) on one MI355X through the Node-API addon.
 *
 *   const s = new AxiHIP(grid, zones, comps, fs, Q0, opt);   // throws when no device
 *   s.step(n); const Q = s.read();   // 5 Float32Array planes; s.src, s.mfx diagnostics
 *   s.setComps(c); s.setZones(z); s.setFs(fs); s.setState(Q); s.finish(); s.dispose();
 *   saveCheckpoint(file, Q, comps) / loadCheckpoint(file, N, 32): the long runs' files.
 * grid = {W, H, xf, rf, ...} (buildCFDGrid, src/cfd_grid.js:28-43; axiGrid below
 * derives xc, rc, dx, dr from the faces), zones = buildZones' typed arrays.
 * ========================================================================== */
const { loadAddon } = require('./hip_solver.js');

/* Grid record from Float64 faces (src/cfd_grid.js:36-42). */
function axiGrid(xf, rf, level) {
  const W = xf.length - 1, H = rf.length - 1;
  const xc = new Float64Array(W), dx = new Float64Array(W), rc = new Float64Array(H), dr = new Float64Array(H);
  for (let i = 0; i < W; i++) { xc[i] = 0.5 * (xf[i] + xf[i + 1]); dx[i] = xf[i + 1] - xf[i]; }
  for (let j = 0; j < H; j++) { rc[j] = 0.5 * (rf[j] + rf[j + 1]); dr[j] = rf[j + 1] - rf[j]; }
  return { W, H, N: W * H, xf: Float64Array.from(xf), rf: Float64Array.from(rf), xc, rc, dx, dr, level };
}

const flat5 = (Q, N) => { const b = new Float64Array(5 * N); for (let q = 0; q < 5; q++) b.set(Q[q], q * N); return b; };

class AxiHIP {
  constructor(grid, zones, comps, fs, Q0, opt) {
    this.kind = 'hip';
    this.opt = Object.assign({ cfl: 0.6, order: 2, visc: true, hf: 0.99 * 43e6 }, opt || {});   // src/cfd_gpu.js:208
    this.grid = grid; this.W = grid.W; this.H = grid.H; this.iter = 0;
    const A = loadAddon();
    this.h = A.axiCreate(Float64Array.from(grid.xf), Float64Array.from(grid.rf), zones, Float64Array.from(comps),
      Float64Array.from([fs.r, fs.u, fs.p]), flat5(Q0, grid.W * grid.H));
    this.z = zones; this.comps = Float64Array.from(comps); this.fs = fs;
    this._opt = null;
    const N = grid.W * grid.H;
    this._buf = new Float32Array(5 * N); this.src = new Float32Array(N); this.mfx = new Float32Array(N);
  }
  _syncOpt() {
    const o = this.opt, key = [o.cfl, o.order, o.visc, o.hf, o.global, o.dtRatio].join(',');
    if (key !== this._opt) {
      loadAddon().axiSetOpt(this.h, o.cfl, o.order, o.visc ? 1 : 0, o.hf || 0, o.global ? 1 : 0, o.dtRatio || 0);
      this._opt = key;
    }
  }
  setComps(c) { this.comps = Float64Array.from(c); loadAddon().axiSetComps(this.h, this.comps); }
  setZones(z) { this.z = z; this._dispZones = false; loadAddon().axiSetZones(this.h, z, this.W * this.H); }
  setFs(fs) { this.fs = fs; loadAddon().axiSetFs(this.h, Float64Array.from([fs.r, fs.u, fs.p])); }
  setState(Q) { loadAddon().axiSetState(this.h, flat5(Q, this.W * this.H)); }
  step(n) { this._syncOpt(); loadAddon().axiStep(this.h, n || 1); this.iter += n || 1; }
  read(Q) {
    const N = this.W * this.H;
    loadAddon().axiRead(this.h, this._buf, this.src, this.mfx);
    if (!Q) Q = [0, 1, 2, 3, 4].map(() => new Float32Array(N));
    for (let q = 0; q < 5; q++) Q[q].set(this._buf.subarray(q * N, (q + 1) * N));
    return Q;
  }
  finish() { loadAddon().axiSync(this.h); }
  /* kernel form of the uncapped steps: 0 face + cell kernels, 1 .. 4 fused tile kernels (the same bits) */
  setFused(form) { loadAddon().axiSetFused(this.h, form); }
  /* keep the fastest kernel form for this grid (state kept): {form, msPerIter} */
  autotune(iters = 50) { this._syncOpt(); return loadAddon().axiAutotune(this.h, iters); }
  dispose() { if (this.h) { loadAddon().axiDispose(this.h); this.h = null; } }
  /* Sections the device monitor sums (at most 4): [{col, rows: Uint8Array(H)}]
   * as cfd_post.js sectionRows builds them. */
  monitorSections(list) {
    const cols = Int32Array.from(list.map(x => x.col)), rows = new Uint8Array(list.length * this.H);
    list.forEach((x, q) => rows.set(x.rows, q * this.H));
    loadAddon().axiMonitorSections(this.h, cols, rows);
    this.nsec = list.length;
  }
  /* The readback of the convergence loop on the device (FP64): thrust parts
   * and blade-row powers of cfdForces with the solver's own sources, the
   * density residual sums against the previous monitor call, and the
   * section sums {mdot, w, wTt}, and the count of non-physical cells of the
   * divergence check (bad). Waits for the GPU; reads back 29 numbers. */
  monitor(comps, P0) {
    const o = this._mon || (this._mon = new Float64Array(29));
    loadAddon().axiMonitor(this.h, comps instanceof Float64Array ? comps : Float64Array.from(comps), P0, o);
    const sec = [];
    for (let q = 0; q < o[15]; q++) sec.push({ mdot: o[16 + 3 * q], w: o[17 + 3 * q], wTt: o[18 + 3 * q] });
    return { F: o[0], Fb: o[1], Fw: o[2], mfuel: o[3], pw: o.slice(4, 12), resSum2: o[12], resN: o[13], resValid: o[14] !== 0, sec,
      bad: o[28] };
  }
  /* Section sums of the current state on the device for [{col, rows}] (any
   * number up to 256): per section a Float64Array(11) of mass flow, w, area,
   * sum p dA and the w-weighted Tt, Pt, T, p, |V|, M, Y (cfd_post.js
   * sectionAverages / stationAverages turn them into averages). */
  sections(list) {
    const n = list.length, cols = Int32Array.from(list.map(x => x.col)), rows = new Uint8Array(n * this.H), o = new Float64Array(11 * n);
    list.forEach((x, q) => rows.set(x.rows, q * this.H));
    loadAddon().axiSections(this.h, cols, rows, o);
    return list.map((x, q) => o.subarray(11 * q, 11 * q + 11));
  }
  /* Display fields of the current state on the device (cfdDerived,
   * src/cfd_post.js:10-40): {W, H, N, reg, mat, T, Tt, P, Pt, V, M, rho, Vx,
   * Vth, Vr, Y, vort, ds} with Float32Array fields. fl = {T0, P0}; disp: cells
   * under half full show as solid. G: a previous result to refill. */
  derived(fl, disp, G) {
    const N = this.W * this.H;
    if (!this._dispZones) {
      if (!this.z.stream || !this.z.mat) throw new Error('AxiHIP.derived: the zones need stream and mat (buildZones)');
      loadAddon().axiSetDisplayZones(this.h, Int8Array.from(this.z.stream), Uint8Array.from(this.z.mat));
      this._dispZones = true;
    }
    if (!G || G.N !== N) {
      G = { W: this.W, H: this.H, N, reg: new Uint8Array(N), mat: new Uint8Array(N), _f: new Float32Array(13 * N) };
      DERIVED_KEYS.forEach((k, c) => { G[k] = G._f.subarray(c * N, (c + 1) * N); });
    }
    loadAddon().axiDerived(this.h, disp ? 1 : 0, fl.T0, fl.P0, G._f, G.reg, G.mat);
    return G;
  }
  /* Grid sequencing on the device (interpState, src/app_cfd.js:321-358): the
   * state of solver `src` (another grid) carried onto this solver's grid and
   * made its current state. fl = {T0, P0} of the fallback at rest. */
  transferFrom(src, fl) { loadAddon().axiTransfer(this.h, src.h, fl.T0, fl.P0); }
}

/* field order of AxiHIP.derived's planes */
const DERIVED_KEYS = ['T', 'Tt', 'P', 'Pt', 'V', 'M', 'rho', 'Vx', 'Vth', 'Vr', 'Y', 'vort', 'ds'];

/* Checkpoint of a long run in the reference's format (save and setup's load,
 * tools/cfdlib.js:16-21): the five state planes as one raw Float64 file and
 * the component constants as JSON beside it (<file>.comps.json). */
function saveCheckpoint(file, Q, comps) {
  const fs = require('fs'), N = Q[0].length, b = new Float64Array(5 * N);
  for (let q = 0; q < 5; q++) b.set(Q[q], q * N);
  fs.writeFileSync(file, Buffer.from(b.buffer));
  fs.writeFileSync(file + '.comps.json', JSON.stringify(Array.from(comps)));
}
/* {Q, comps} of a checkpoint; a part that is missing or sized for another
 * grid / constants layout comes back null (the reference's loader then keeps
 * the setup's). */
function loadCheckpoint(file, N, ncomps) {
  const fs = require('fs');
  let Q = null, comps = null;
  if (fs.existsSync(file)) {
    const buf = fs.readFileSync(file), b = new Float64Array(buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.byteLength));
    if (b.length === 5 * N) Q = [0, 1, 2, 3, 4].map(q => b.slice(q * N, (q + 1) * N));
  }
  if (fs.existsSync(file + '.comps.json')) {
    const cj = JSON.parse(fs.readFileSync(file + '.comps.json', 'utf8'));
    if (cj.length === ncomps) comps = Float64Array.from(cj);
  }
  return { Q, comps };
}

module.exports = { AxiHIP, axiGrid, saveCheckpoint, loadCheckpoint, DERIVED_KEYS };
