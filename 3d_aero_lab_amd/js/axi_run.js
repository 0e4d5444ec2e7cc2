'use strict';
/* ============================================================================
 * axi_run.js — the reference's long axisymmetric run (tools/longrun.js:2-14)
 * on AxiHIP: n iterations in chunks of 250 with the exit-temperature control
 * after every chunk, thrust and station averages at the end. Every readback
 * is the device's (AxiHIP.monitor / sections, FP64 sums): only the numbers
 * come back, the state once at the end (for a checkpoint).
 *
 *   node axi_run.js <case dir> <iters> [report every] [spec] [model=<file>] [load=<file>] [save=<file>]
 * <case dir> holds the solver inputs (grid faces xf / rf, zones z_*, comps,
 * initial state Q0_*, freestream, control fields, engine description) as
 * written by tests/golden/make_golden_axi.js from the reference's setup; with
 * `spec` the grid, zones and constants are built from the engine description
 * (cfd_grid.js) instead of read, and with model= (a JSON field model
 * description, field_model.js) the initial state is the device's
 * initialState instead of read; load= resumes from a checkpoint of the
 * reference's long-run format, save= writes one at the end.
 * Prints one JSON line: {hist, thrust, stations, ms_per_iter_device_wall}.
 * ========================================================================== */
const fs = require('fs'), path = require('path'), zlib = require('zlib');
const P = require('./cfd_post.js');
const { AxiHIP, axiGrid, saveCheckpoint, loadCheckpoint } = require('./cfd_axi.js');

/* Solver inputs of a case directory. With fromSpec, the grid, zones and
 * component constants are built by cfd_grid.js from the case's engine
 * description (`axispec`) and only the initial state is read. */
function loadCase(dir, fromSpec, modelFile) {
  const m = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
  const T = { f4: Float32Array, f8: Float64Array, u1: Uint8Array, i1: Int8Array };
  const arr = n => { const a = m.arrays[n]; const b = zlib.gunzipSync(fs.readFileSync(path.join(dir, a.file))); return new T[a.dtype](b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
  if (fromSpec) {
    const { grid, zones, comps } = require('./cfd_grid.js').axiSetup(m.scalars.axispec);
    let Q0;
    if (modelFile) {                              // initialState from the engine's field model, on the device
      const FM = require('./field_model.js');
      const fm = new FM.FieldHIP(JSON.parse(fs.readFileSync(modelFile, 'utf8')));
      Q0 = FM.initialState(grid, zones, fm, false);
      fm.dispose();
    } else Q0 = [0, 1, 2, 3, 4].map(q => arr(`Q0_${q}`));
    return { s: m.scalars, grid, zones, comps, Q0 };
  }
  const Q0 = [0, 1, 2, 3, 4].map(q => arr(`Q0_${q}`));
  const zones = {};
  for (const k of ['solid', 'phi', 'ae', 'an', 'comp', 'wgt', 'lmix', 'tgt', 'rate', 'yrate', 'mfd', 'stream']) zones[k] = arr('z_' + k);
  return { s: m.scalars, grid: axiGrid(arr('xf'), arr('rf')), zones, comps: Float64Array.from(arr('comps')), Q0 };
}

function run(dir, iters, every, fromSpec, ckpt, modelFile) {
  const C = loadCase(dir, fromSpec, modelFile), s = C.s;
  let comps = C.comps, Q0 = C.Q0;
  if (ckpt && ckpt.load) {                       // resume (tools/cfdlib.js:16-17)
    const L = loadCheckpoint(ckpt.load, C.grid.N, comps.length);
    if (L.Q) Q0 = L.Q;
    if (L.comps) comps = L.comps;
  }
  const S = new AxiHIP(C.grid, C.zones, comps, s.fs, Q0, Object.assign({}, s.opt, { hf: s.hf }));
  const ctlSt = P.controlStations(s.ctl);
  S.monitorSections(ctlSt.map(st => P.sectionRows(C.grid, C.zones, st.x, st.si)));
  const hist = [];
  let it = 0, f = null;
  const t0 = process.hrtime.bigint();
  let tStep = 0n;
  while (it < iters) {
    const n = Math.min(250, iters - it);
    const a = process.hrtime.bigint();
    S.step(n); S.finish();
    tStep += process.hrtime.bigint() - a;
    it += n;
    // the readback on the device: thrust, blade-row powers and the control sections (FP64)
    f = S.monitor(comps, s.ctl.P0);
    P.controlFromSections(s.type || 'tj', s.ctl, st => P.sectionFromSums(f.sec[st.id === 4 ? 0 : 1]), comps, f.pw, s.ctl.etam, 250);
    S.setComps(comps);
    if (it % every === 0 || it >= iters) hist.push({ it, F: f.F, Fb: f.Fb, Fw: f.Fw, mfuel: f.mfuel, K4: comps[17] });
  }
  const wall = Number(process.hrtime.bigint() - t0) / 1e9;
  const st = P.stations(S, C.grid, C.zones, s.ctl.order, s.type || 'tj');
  const Q = S.read();
  S.dispose();
  if (ckpt && ckpt.save) saveCheckpoint(ckpt.save, Q, comps);
  return { cells: C.grid.N, iters, hist, thrust: { F: f.F, Fb: f.Fb, Fw: f.Fw, mfuel: f.mfuel }, F1D: s.ctl.F1D, stations: st,
    wall_s: wall, step_s: Number(tStep) / 1e9 };
}

if (require.main === module) {
  const pos = process.argv.slice(2).filter(a => !a.includes('=')), kv = {};
  for (const a of process.argv.slice(2).filter(a => a.includes('='))) { const [k, v] = a.split('='); kv[k] = v; }
  const [dir, iters, every, setup] = pos;
  console.log(JSON.stringify(run(dir, +iters, +(every || 2000), setup === 'spec', { load: kv.load, save: kv.save }, kv.model)));
}
module.exports = { loadCase, run };
