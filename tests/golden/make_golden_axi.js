'use strict';
/* ============================================================================
 * make_golden_axi.js — golden vectors of the reference's axisymmetric engine
 * solver (CPUSolver, src/cfd_cpu.js:50-210) and its post-processing
 * (cfdForces / cfdSection / cfdControl, src/cfd_post.js:44-123), by running the
 * reference's JavaScript (read-only checkout, default /root/reference, override
 * with AERO_REF) under Node.
 *
 *   node tests/golden/make_golden_axi.js [case ...]
 * cases: xcoarse medium longrun (TJ), hbtf_xcoarse hbtf_medium lbtf_xcoarse,
 * longrun_hbtf longrun_lbtf (the README.md:71-75 validation rows), setups,
 * interp, tj_coarse, checkpoint, options, field (the field model: samples and
 * initial states)
 *
 * The reference's own harness (tools/cfdlib.js:6-20) uses `??`, which this
 * Node (v12) cannot parse; the same setup sequence is written out here:
 * solveEngine -> buildSpec -> FieldModel -> buildCFDGrid -> buildZones ->
 * compParams -> initialState -> CPUSolver (cfl 0.6, order 2, visc, hf = etb LHV),
 * with Tt7 2100 K for the LBTF and 2000 K otherwise (tools/cfdlib.js:8).
 * Outputs (data only): grid faces, zones, comps, freestream, initial state,
 * states / diagnostics after N iterations, thrust and control values.
 * ========================================================================== */
const fs = require('fs');
const path = require('path');
const vm = require('vm');
const zlib = require('zlib');
const crypto = require('crypto');

const REF = process.env.AERO_REF || '/root/reference';
const code = ['physics', 'geometry', 'field', 'cfd_grid', 'cfd_cpu', 'cfd_post']
  .map(f => fs.readFileSync(path.join(REF, 'src', f + '.js'), 'utf8').replace(/^if \(typeof module[^\n]*\n/gm, '')).join('\n');
vm.runInThisContext(code + '\nglobalThis.AXI={cfdDerived,cfdControl,solveEngine,ENGINE_TYPES,DEFAULT_EFF,buildSpec,FieldModel,buildCFDGrid,buildZones,compParams,initialState,CPUSolver,cfdForces,cfdSection,spoolRatios,RGAS,GAS};');
const A = globalThis.AXI;
const OUT = __dirname;
const { fieldDesc } = require('./field_desc.js');

function setup(type, level) {
  const T = A.ENGINE_TYPES[type];
  const P = { type, design: Object.assign({}, T.design), eff: Object.assign({}, A.DEFAULT_EFF), alt: 0, M0: 0.25, thr: 1, ab: false, Tt7: type === 'lbtf' ? 2100 : 2000, nozzleType: T.nozzle };
  const sol = A.solveEngine(P), op = sol.op, fl = sol.fl, nz = op.noz;
  const spec = A.buildSpec(type, 1, { r8: Math.sqrt(nz.Athroat / Math.PI), r9: Math.sqrt(Math.max(nz.A, nz.Athroat) / Math.PI), A9: sol.ref.noz && sol.ref.noz.A, A19: sol.ref.nozB && sol.ref.nozB.A });
  const model = new A.FieldModel(spec, sol, P);
  const grid = A.buildCFDGrid(spec, level);
  const zones = A.buildZones(grid, spec, model, op, P);
  const comps = A.compParams(type, op, P);
  const Q0 = A.initialState(grid, zones, spec, model, fl, false);
  const fsv = { r: fl.P0 / (A.RGAS * fl.T0), u: fl.V0, p: fl.P0 };
  const hf = P.eff.etb * P.eff.LHV;
  const S = new A.CPUSolver(grid, zones, comps, fsv, Q0, { cfl: 0.6, order: 2, visc: true, hf });
  return { P, sol, op, fl, spec, grid, zones, comps, base: Float64Array.from(comps), Q0, fsv, hf, S, type, m: model };
}

function sha(buf) { return crypto.createHash('sha256').update(buf).digest('hex'); }
class Writer {
  constructor(name) { this.dir = path.join(OUT, name); fs.mkdirSync(this.dir, { recursive: true }); this.m = { case: name, arrays: {}, scalars: {} }; }
  arr(name, ta, hashOnly) {
    const buf = Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength);
    const dtype = ta instanceof Float32Array ? 'f4' : ta instanceof Float64Array ? 'f8' : ta instanceof Uint8Array ? 'u1' : ta instanceof Int8Array ? 'i1' : '?';
    let file = null;
    if (!hashOnly) { file = name + '.bin.gz'; fs.writeFileSync(path.join(this.dir, file), zlib.gzipSync(buf, { level: 9 })); }
    this.m.arrays[name] = { file, dtype, n: ta.length, sha256: sha(buf) };
  }
  set(k, v) { this.m.scalars[k] = v; }
  close() { fs.writeFileSync(path.join(this.dir, 'manifest.json'), JSON.stringify(this.m, null, 1)); console.log(`wrote ${this.dir}`); }
}
const ZONE_KEYS = ['solid', 'phi', 'ae', 'an', 'comp', 'wgt', 'lmix', 'tgt', 'rate', 'yrate', 'mfd', 'stream', 'mat'];

/* The engine geometry and 1D operating point the axisymmetric setup reads
 * (buildCFDGrid / buildZones / compParams, src/cfd_grid.js:28-276), as data:
 * stream walls and liners as their polyline points, body polygons, blade rows,
 * combustor / afterburner extents, exits, cycle temperatures and mass flows. */
function axiSpec(C) {
  const sp = C.spec, op = C.op, lp = L => L.pts.map(p => [p[0], p[1]]);
  const streams = {};
  for (const k of sp.order) { const st = sp.streams[k]; streams[k] = { x0: st.x0, x1: st.x1, lo: lp(st.lo), hi: lp(st.hi) }; }
  const opk = {};
  for (const k of ['Tt2', 'Tt3', 'Tt4', 'Tt5', 'Tt6', 'Tt7', 'Tt13', 'Tt25', 'Tt45', 'abOn', 'mdot', 'f', 'mcore', 'mbyp', 'm7'])
    if (op[k] !== undefined && op[k] !== null && !(typeof op[k] === 'number' && isNaN(op[k]))) opk[k] = op[k];
  return {
    type: sp.type, s: sp.s, level: C.grid.level, domain: { xb: sp.domain.xb, rb: sp.domain.rb }, view: { len: sp.view.len, Rmax: sp.view.Rmax },
    order: sp.order, streams,
    bodies: sp.bodies.map(b => ({ key: b.key, mat: b.mat, poly: b.poly })),
    rows: sp.rows.map(r => ({ x0: r.x0, x1: r.x1, stream: r.stream, rotor: !!r.rotor, comp: r.comp })),
    comb: { xd: sp.comb.xd, x1: sp.comb.x1, stream: sp.comb.stream, linerIn: lp(sp.comb.linerIn), linerOut: lp(sp.comb.linerOut) },
    ab: sp.ab ? { xFH: sp.ab.xFH, x1: sp.ab.x1, stream: sp.ab.stream, rings: sp.ab.rings } : null,
    exitsGeom: sp.exitsGeom.map(e => ({ x: e.x, R: e.R })),
    op: opk, eff: { ef: C.P.eff.ef, ec: C.P.eff.ec, et: C.P.eff.et },
  };
}

function dumpSetup(w, C, q0hash) {
  const g = C.grid;
  w.set('axispec', axiSpec(C));
  w.arr('xf', g.xf); w.arr('rf', g.rf);
  for (const k of ZONE_KEYS) w.arr('z_' + k, C.zones[k]);
  w.arr('comps', C.comps);
  for (let q = 0; q < 5; q++) w.arr('Q0_' + q, C.Q0[q], q0hash);
  w.set('W', g.W); w.set('H', g.H); w.set('N', g.N); w.set('type', C.type);
  w.set('fs', C.fsv); w.set('hf', C.hf); w.set('opt', { cfl: 0.6, order: 2, visc: true });
  w.set('fl', { T0: C.fl.T0, P0: C.fl.P0 });
  // what cfdControl reads from spec / op (src/cfd_post.js:110-123)
  w.set('ctl', { s: C.spec.s, comb: { x1: C.spec.comb.x1, stream: C.spec.comb.stream }, ab: C.spec.ab ? { x1: C.spec.ab.x1, stream: C.spec.ab.stream } : null,
    order: C.spec.order, Tt4: C.op.Tt4, Tt7: C.op.Tt7 || null, abOn: !!C.op.abOn, etam: C.P.eff.etam, P0: C.fl.P0, F1D: C.op.F });
}
function dumpState(w, tag, C, hashOnly) {
  const S = C.S;
  for (let q = 0; q < 5; q++) w.arr(`${tag}_Q${q}`, S.Q[q], hashOnly);
  w.arr(`${tag}_src`, S.src, hashOnly); w.arr(`${tag}_mfx`, S.mfx, hashOnly); w.arr(`${tag}_dt`, S.dt, hashOnly);
}
function thrust(C) {
  const f = A.cfdForces(C.grid, C.zones, C.S.Q, C.comps, C.fl.P0, C.hf, { S: C.S.src, F: C.S.mfx });
  return { F: f.F, Fb: f.Fb, Fw: f.Fw, mfuel: f.mfuel, pw: Array.from(f.pw) };
}

function caseXcoarse() {
  const w = new Writer('axi_tj_xcoarse');
  const C = setup('tj', 'xcoarse');
  dumpSetup(w, C, false);
  const t0 = Date.now();
  let it = 0;
  const hist = [];
  for (const target of [1, 10, 100, 250, 500]) {
    C.S.step(target - it); it = target;
    dumpState(w, 'it' + it, C, !(it === 10 || it === 100 || it === 500));
    hist.push({ it, ...thrust(C) });
  }
  w.set('ms_per_iter', (Date.now() - t0) / it);
  w.set('hist', hist);
  derived(w, C);
  controlAndSection(w, C);
  w.close();
}

/* display fields (cfdDerived, src/cfd_post.js:10-40) of the solver's current
 * state, with and without the display solid rule */
const DER_KEYS = ['reg', 'mat', 'T', 'Tt', 'P', 'Pt', 'V', 'M', 'rho', 'Vx', 'Vth', 'Vr', 'Y', 'vort', 'ds'];
function derived(w, C) {
  for (const disp of [false, true]) {
    const z = Object.assign({}, C.zones);   // cfdDerived caches its display mask on the zones record
    const G = A.cfdDerived(C.grid, z, C.S.Q, C.fl, null, disp);
    for (const k of DER_KEYS) w.arr(`der${disp ? 1 : 0}_${k}`, G[k]);
  }
}

/* one control step on the last state (the long-run loop calls it every 250
 * iterations) and the combustor-exit section average */
function controlAndSection(w, C) {
  const f0 = A.cfdForces(C.grid, C.zones, C.S.Q, C.comps, C.fl.P0, 0, { S: C.S.src, F: C.S.mfx });
  const ctl = A.cfdControl(C.type, C.spec, C.grid, C.zones, C.S.Q, C.S.mfx, C.comps, C.base, f0.pw, C.op, C.P.eff.etam, 250);
  w.set('control', { Tt4: ctl.Tt4, Tt7: ctl.Tt7, ratios: ctl.ratios, comps_after: Array.from(C.comps) });
  const sec = A.cfdSection(C.grid, C.zones, C.S.Q, C.spec.comb.x1 - 0.004 * C.spec.s, C.spec.order.indexOf(C.spec.comb.stream), C.S.mfx);
  w.set('section', sec);
}

function caseMedium() {
  const w = new Writer('axi_tj_medium');
  const C = setup('tj', 'medium');
  dumpSetup(w, C, false);
  const t0 = Date.now();
  C.S.step(2);
  w.set('ms_per_iter', (Date.now() - t0) / 2);
  dumpState(w, 'it2', C, true);
  w.set('hist', [{ it: 2, ...thrust(C) }]);
  w.close();
}

/* README.md:69-75 runs: FRESH=1 node tools/longrun.js <type> <level> <iters> 2000 (control every 250 iterations). */
function caseLongrun(type, level, iters0, name) {
  const iters = +(process.env.AXI_LONG_ITERS || iters0), every = 2000;
  const w = new Writer(name);
  const C = setup(type, level);
  const hist = [];
  const t0 = Date.now();
  for (let n = 0; n < iters; n += every) {
    for (let m = 0; m < every; m += 250) {
      C.S.step(250);
      const f0 = A.cfdForces(C.grid, C.zones, C.S.Q, C.comps, C.fl.P0, 0, { S: C.S.src, F: C.S.mfx });
      A.cfdControl(C.type, C.spec, C.grid, C.zones, C.S.Q, C.S.mfx, C.comps, C.base, f0.pw, C.op, C.P.eff.etam, 250);
    }
    const h = { it: n + every, ...thrust(C), K4: C.comps[17], s: (Date.now() - t0) / 1000 };
    hist.push(h);
    console.log(JSON.stringify(h));
    w.set('hist', hist);
    w.m.arrays = {};
    for (let q = 0; q < 5; q++) w.arr(`Q${q}`, C.S.Q[q], true);
    w.arr('comps', C.comps);
    w.close();
  }
}

/* setup, then full states after 10 and 100 iterations (thrust history at 1, 10, 100) */
function caseShort(type, level) {
  const w = new Writer(`axi_${type}_${level}`);
  const C = setup(type, level);
  dumpSetup(w, C, false);
  const t0 = Date.now();
  let it = 0;
  const hist = [];
  for (const target of [1, 10, 100]) {
    C.S.step(target - it); it = target;
    dumpState(w, 'it' + it, C, it === 1);
    hist.push({ it, ...thrust(C) });
  }
  w.set('ms_per_iter', (Date.now() - t0) / it);
  w.set('hist', hist);
  derived(w, C);
  controlAndSection(w, C);
  w.close();
}
function caseSetupOnly(type, level) {
  const w = new Writer(`axi_${type}_${level}`);
  const C = setup(type, level);
  dumpSetup(w, C, false);
  C.S.step(2);
  dumpState(w, 'it2', C, true);
  w.set('hist', [{ it: 2, ...thrust(C) }]);
  w.close();
}

/* setups only (engine description + sha256 of every setup array) of the
 * three engines at every grid level: the setup restatement's fixtures */
function caseSetups() {
  const out = {};
  for (const type of ['tj', 'hbtf', 'lbtf']) for (const level of ['xcoarse', 'coarse', 'medium', 'fine']) {
    const C = setup(type, level), z = C.zones, h = {};
    h.xf = sha(Buffer.from(C.grid.xf.buffer)); h.rf = sha(Buffer.from(C.grid.rf.buffer)); h.comps = sha(Buffer.from(C.comps.buffer));
    for (const k of ZONE_KEYS) h['z_' + k] = sha(Buffer.from(z[k].buffer, z[k].byteOffset, z[k].byteLength));
    out[`${type}_${level}`] = { W: C.grid.W, H: C.grid.H, axispec: axiSpec(C), sha256: h };
  }
  fs.mkdirSync(path.join(OUT, 'axi_setups'), { recursive: true });
  fs.writeFileSync(path.join(OUT, 'axi_setups', 'manifest.json'), JSON.stringify({ case: 'axi_setups', setups: out }));
  console.log('wrote axi_setups');
}

/* interpState (grid sequencing, src/app_cfd.js:331-358 with fracIdx :321-326),
 * taken as text from the reference's app module and evaluated here (the rest
 * of that module needs a browser page): the xcoarse reference state carried
 * onto the medium grid for the TJ and the HBTF, as sha256 of the five planes
 * (Float64, and rounded to Float32 as a solver's state holds them). */
function refInterp() {
  const src = fs.readFileSync(path.join(REF, 'src', 'app_cfd.js'), 'utf8');
  const grab = name => {
    const i = src.indexOf(`function ${name}(`);
    let d = 0, k = src.indexOf('{', i);
    for (; k < src.length; k++) { if (src[k] === '{') d++; else if (src[k] === '}' && --d === 0) break; }
    return src.slice(i, k + 1);
  };
  vm.runInThisContext(grab('fracIdx') + '\n' + grab('interpState') + '\nglobalThis.AXI_INTERP = { fracIdx, interpState };');
  return globalThis.AXI_INTERP;
}
function caseInterp() {
  const I = refInterp(), out = {};
  for (const [type, iters] of [['tj', 500], ['hbtf', 100]]) {
    const A0 = setup(type, 'xcoarse'), B = setup(type, 'medium');
    A0.S.step(iters);
    globalThis.CUR = { sol: { fl: B.fl } };      // the all-solid fallback reads the freestream from the app state
    const Q = I.interpState(A0.grid, A0.zones, A0.S.Q, B.grid, B.zones);
    out[type] = { from: 'xcoarse', to: 'medium', iters, sha256: Q.map(a => sha(Buffer.from(a.buffer))),
      sha256_f32: Q.map(a => sha(Buffer.from(Float32Array.from(a).buffer))),
      sample: Q.map(a => Array.from(a.subarray(0, 64))) };
  }
  fs.mkdirSync(path.join(OUT, 'axi_interp'), { recursive: true });
  fs.writeFileSync(path.join(OUT, 'axi_interp', 'manifest.json'), JSON.stringify({ case: 'axi_interp', interp: out }));
  console.log('wrote axi_interp');
}

/* Checkpoint format of the reference's long runs (save, tools/cfdlib.js:21):
 * the state after 10 TJ xcoarse iterations written by that function (taken
 * as text, like interpState) to a scratch file; its bytes go in as sha256. */
function caseCheckpoint() {
  const src = fs.readFileSync(path.join(REF, 'tools', 'cfdlib.js'), 'utf8');
  const i = src.indexOf('function save(');
  const j = src.indexOf('\n', i);
  globalThis.fs = fs;                      // the function writes with the file module
  vm.runInThisContext(src.slice(i, j) + '\nglobalThis.AXI_SAVE = save;');
  const C = setup('tj', 'xcoarse');
  C.S.step(10);
  const os = require('os'), f = path.join(os.tmpdir(), 'axi_ckpt_ref.bin');
  globalThis.AXI_SAVE(C, f);
  const bin = fs.readFileSync(f), js = fs.readFileSync(f + '.comps.json', 'utf8');
  fs.mkdirSync(path.join(OUT, 'axi_ckpt'), { recursive: true });
  fs.writeFileSync(path.join(OUT, 'axi_ckpt', 'manifest.json'), JSON.stringify({ case: 'axi_ckpt', state: 'axi_tj_xcoarse it10',
    bin_bytes: bin.length, bin_sha256: sha(bin), comps_json: js }));
  fs.unlinkSync(f); fs.unlinkSync(f + '.comps.json');
  console.log('wrote axi_ckpt');
}

/* CPUSolver's option paths (src/cfd_cpu.js:53, 193-203) on TJ xcoarse: the
 * global and dtRatio dt caps and first order without the viscous r fluxes,
 * 10 iterations each; and dtRatio 3 for 30 iterations, then 5 for 30 more.
 * The conserved planes go in full (Float64). */
function caseOptions() {
  const w = new Writer('axi_tj_options');
  const dumpQ = (tag, C) => { for (let q = 0; q < 5; q++) w.arr(`${tag}_Q${q}`, C.S.Q[q]); };
  for (const [tag, o] of [['global', { global: true }], ['dtr3', { dtRatio: 3 }], ['o1e', { order: 1, visc: false }]]) {
    const C = setup('tj', 'xcoarse');
    Object.assign(C.S.opt, o);
    C.S.step(10);
    dumpQ(`${tag}_it10`, C);
    if (tag === 'global') w.arr('global_it10_dt', C.S.dt);
  }
  const C = setup('tj', 'xcoarse');
  C.S.opt.dtRatio = 3; C.S.step(30);
  C.S.opt.dtRatio = 5; C.S.step(30);
  dumpQ('dtr35_it60', C);
  w.set('runs', { global: { global: true, iters: 10 }, dtr3: { dtRatio: 3, iters: 10 }, o1e: { order: 1, visc: false, iters: 10 },
    dtr35: { dtRatio: [3, 5], iters: [30, 30] } });
  w.close();
}

/* The engine field model as data and its outputs: the tables that the
 * FieldModel constructor derives from the 1D cycle (src/field.js:255-294:
 * stream knots after assignMach, processed exits, inlet face velocity, hot
 * gas) with the geometry it samples (stream walls, bodies, combustor, mixer,
 * afterburner, lip, outer line), FieldModel.sample (:318-488) at seeded
 * points, and initialState (src/cfd_grid.js:279-305) on all twelve setups. */
function fieldModelOf(type, ab) {
  const T = A.ENGINE_TYPES[type];
  const P = { type, design: Object.assign({}, T.design), eff: Object.assign({}, A.DEFAULT_EFF), alt: 0, M0: 0.25, thr: 1, ab: !!ab, Tt7: type === 'lbtf' ? 2100 : 2000, nozzleType: T.nozzle };
  const sol = A.solveEngine(P), nz = sol.op.noz;
  const spec = A.buildSpec(type, 1, { r8: Math.sqrt(nz.Athroat / Math.PI), r9: Math.sqrt(Math.max(nz.A, nz.Athroat) / Math.PI), A9: sol.ref.noz && sol.ref.noz.A, A19: sol.ref.nozB && sol.ref.nozB.A });
  return { spec, sol, P, m: new A.FieldModel(spec, sol, P) };
}
function mulberry32(a) {
  return function () { a |= 0; a = a + 0x6D2B79F5 | 0; let t = Math.imul(a ^ a >>> 15, 1 | a); t = t + Math.imul(t ^ t >>> 7, 61 | t) ^ t; return ((t ^ t >>> 14) >>> 0) / 4294967296; };
}
const SAMPLE_KEYS = ['T', 'P', 'Tt', 'Pt', 'V', 'Vx', 'Vth', 'M', 'a', 'rho', 'g', 'cp', 'R'];
function caseField() {
  const w = new Writer('axi_field');
  const models = {};
  for (const [name, type, ab] of [['tj', 'tj', false], ['hbtf', 'hbtf', false], ['lbtf', 'lbtf', false], ['tj_ab', 'tj', true], ['lbtf_ab', 'lbtf', true]]) {
    const F = fieldModelOf(type, ab), sp = F.spec, m = F.m, d = fieldDesc(sp, m, A.GAS.cold);
    models[name] = d;
    // seeded points: a third over the engine, a third inside the streams (every
    // wall distance), a third over the whole domain (jets, plume, shock cells)
    const rnd = mulberry32(20261205 + name.length * 7 + (ab ? 1 : 0)), n = 3000;
    const X = new Float64Array(n), R = new Float64Array(n), TH = new Float64Array(n), out = new Float64Array(n * 14);
    const xb = sp.domain.xb, rb = sp.domain.rb, len = sp.view.len, Rm = sp.view.Rmax;
    for (let p = 0; p < n; p++) {
      const kind = p % 3;
      let x, r;
      if (kind === 0) { x = -0.3 * sp.s + rnd() * (len + 0.6 * sp.s); r = rnd() * 1.15 * Rm; }
      else if (kind === 1) { const st = sp.streams[sp.order[Math.floor(rnd() * sp.order.length)]]; x = st.x0 + rnd() * (st.x1 - st.x0); const lo = st.lo.at(x), hi = st.hi.at(x); r = lo + rnd() * (hi - lo); }
      else { x = xb[0] + rnd() * (xb[xb.length - 1] - xb[0]); r = rnd() * rb[rb.length - 1]; }
      const th = rnd() < 0.5 ? 0 : rnd() * 2 * Math.PI;
      X[p] = x; R[p] = r; TH[p] = th;
      const o = m.sample(x, r, th, {});
      out[p * 14] = o.solid ? 1 : o.region === 'ext' ? 0 : 2 + sp.order.indexOf(o.stream);
      for (let q = 0; q < 13; q++) out[p * 14 + 1 + q] = o.solid ? NaN : o[SAMPLE_KEYS[q]];
    }
    w.arr(`${name}_x`, X); w.arr(`${name}_r`, R); w.arr(`${name}_th`, TH); w.arr(`${name}_out`, out);
  }
  w.set('models', models);
  // initialState on the twelve setups (the model at the setup's operating point)
  const q0 = {};
  for (const type of ['tj', 'hbtf', 'lbtf']) for (const level of ['xcoarse', 'coarse', 'medium', 'fine']) {
    const C = setup(type, level), N = C.grid.N, stride = 17, ns = Math.ceil(N / stride);
    const smp = new Float64Array(5 * ns);
    for (let q = 0; q < 5; q++) for (let k = 0; k < ns; k++) smp[q * ns + k] = C.Q0[q][k * stride];
    w.arr(`q0s_${type}_${level}`, smp);
    q0[`${type}_${level}`] = { stride, sha256: C.Q0.map(a => sha(Buffer.from(a.buffer))), sha256_f32: C.Q0.map(a => sha(Buffer.from(Float32Array.from(a).buffer))),
      uniform_sha256: A.initialState(C.grid, C.zones, C.spec, C.m, C.fl, true).map(a => sha(Buffer.from(a.buffer))) };
  }
  w.set('initial_state', q0);
  w.close();
}

const want = process.argv.slice(2);
const all = {
  xcoarse: caseXcoarse, medium: caseMedium,
  longrun: () => caseLongrun('tj', 'medium', 14000, 'axi_tj_medium_longrun'),
  hbtf_xcoarse: () => caseShort('hbtf', 'xcoarse'), lbtf_xcoarse: () => caseShort('lbtf', 'xcoarse'),
  hbtf_medium: () => caseSetupOnly('hbtf', 'medium'), setups: caseSetups, interp: caseInterp,
  tj_coarse: () => caseSetupOnly('tj', 'coarse'), checkpoint: caseCheckpoint,
  longrun_hbtf: () => caseLongrun('hbtf', 'medium', 14000, 'axi_hbtf_medium_longrun'),
  longrun_lbtf: () => caseLongrun('lbtf', 'xcoarse', 10000, 'axi_lbtf_xcoarse_longrun'),
  options: caseOptions, field: caseField,
};
for (const k of (want.length ? want : ['xcoarse', 'medium'])) { const t0 = Date.now(); all[k](); console.log(`${k}: ${((Date.now() - t0) / 1000).toFixed(1)} s`); }
