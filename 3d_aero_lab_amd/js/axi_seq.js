'use strict';
/* ============================================================================
 * axi_seq.js — the reference application's convergence loop for the
 * axisymmetric solver on AxiHIP (setupCFD / cfdRead, src/app_cfd.js:270-297,
 * 361-404), headless: read back every `every` iterations, density residual,
 * exit-temperature control once past 1500 iterations (and 300 since the
 * start of the current grid), convergence when the residual is below 2e-5 or
 * the thrust has stayed within 0.3 % for 1800 iterations with the residual
 * below 6e-5. With grid sequencing, the run starts on the coarse grid and
 * moves to the target grid (the state carried over on the device,
 * AxiHIP.transferFrom = interpState; control gains carried) once the coarse
 * grid has converged after 2500 iterations, or after 9000.
 *
 *   node axi_seq.js <coarse case dir> <target case dir> [every] [cap]   (sequenced)
 *   node axi_seq.js - <target case dir> [every] [cap]                   (target grid only)
 * Case directories as written by tests/golden/make_golden_axi.js (setup arrays
 * and initial state). Prints one JSON line: per-grid iterations, residual and
 * thrust histories, final thrust, wall time.
 * Each readback is the device monitor (AxiHIP.monitor: thrust, blade-row
 * powers, residual sums, the control sections and the divergence count in
 * FP64, 29 numbers back) where the application reads the whole state back.
 * Divergence (src/app_cfd.js:377-378, 405-409): a readback that finds a cell
 * k = 0, 7, 14, ... with density or energy not positive abandons the run; it
 * starts again from the initial state and the setup's constants (the whole
 * sequence when sequenced) with the CFL lowered by 0.2 (not below 0.3) and
 * first-order reconstruction, at most `maxRestarts` times.
 * ========================================================================== */
const P = require('./cfd_post.js');
const { AxiHIP } = require('./cfd_axi.js');
const { loadCase } = require('./axi_run.js');

function makeAxiHIP(C, comps, Q0, opt) {
  return new AxiHIP(C.grid, C.zones, comps, C.s.fs, Q0, Object.assign({}, C.s.opt, { hf: C.s.hf }, opt));
}

/* Runs one grid until the reference's criteria (from the state of solver
 * `from` when given, carried onto this grid). Returns the solver (the caller
 * disposes it), or null when a readback found non-physical cells. */
function runGrid(C, comps, from, seq, every, cap, log, opt, make) {
  const s = C.s;
  const S = make(C, comps, C.Q0, opt);
  if (from) S.transferFrom(from, s.fl);
  const minIt = seq ? 2500 : 3000;
  let it = 0, res = NaN, f = null;
  const lastF = [];
  S.monitorSections(P.controlStations(s.ctl).map(st => P.sectionRows(C.grid, C.zones, st.x, st.si)));
  const t0 = process.hrtime.bigint();
  for (;;) {
    S.step(every); it += every;
    const gate = log.itTotal + it > 1500 && it > 300;
    // with the solver's sources the forces do not depend on the heating value
    const m = S.monitor(comps, s.ctl.P0);
    if (m.resValid) res = Math.sqrt(m.resSum2 / Math.max(1, m.resN)) / every;
    if (m.bad) { S.dispose(); log.itTotal += it; return null; }
    f = m;
    if (gate) {
      P.controlFromSections(s.type || 'tj', s.ctl, st => P.sectionFromSums(m.sec[st.id === 4 ? 0 : 1]), comps, m.pw, s.ctl.etam, every);
      S.setComps(comps);
    }
    log.hist.push({ grid: C.grid.W + 'x' + C.grid.H, it: log.itTotal + it, res, F: f.F });
    lastF.push([it, f.F]);
    while (lastF.length > 2 && lastF[1][0] < it - 2000) lastF.shift();
    const span = lastF.length > 1 ? it - lastF[0][0] : 0;
    let fmin = Infinity, fmax = -Infinity;
    for (const [, v] of lastF) { fmin = Math.min(fmin, v); fmax = Math.max(fmax, v); }
    const steady = span >= 1800 && (fmax - fmin) < 0.003 * Math.max(1, Math.abs(f.F)) && res < 6e-5;
    if (it > minIt && (res < 2e-5 || steady || (!seq && it > cap))) break;
    if (seq && it > 9000) break;
    if (!seq && it >= cap) break;
  }
  const wall = Number(process.hrtime.bigint() - t0) / 1e9;
  log.itTotal += it;
  log.grids.push({ W: C.grid.W, H: C.grid.H, iters: it, res, F: f.F, wall_s: wall });
  return S;
}

/* One attempt of the whole run (coarse grid first when sequenced); null on divergence. */
function attempt(coarseDir, T, every, cap, opt, make, load, log) {
  const comps = Float64Array.from(T.comps);
  let from = null;
  if (coarseDir && coarseDir !== '-') {
    const Cc = load(coarseDir), cc = Float64Array.from(Cc.comps);
    from = runGrid(Cc, cc, null, true, every, cap, log, opt, make);
    if (!from) return null;
    comps[17] = cc[17]; comps[29] = cc[29];                 // control gains carried (src/app_cfd.js:281)
  }
  const S = runGrid(T, comps, from, false, every, cap, log, opt, make);
  if (from) from.dispose();
  if (!S) return null;
  const Q = S.read();                                       // the final state, once
  S.dispose();
  return Q;
}

/* o: {makeSolver(C, comps, Q0, opt), loadCase(dir), maxRestarts} (defaults: AxiHIP, the fixture loader, 3). */
function run(coarseDir, targetDir, every, cap, o = {}) {
  const make = o.makeSolver || makeAxiHIP, load = o.loadCase || loadCase, maxRestarts = o.maxRestarts === undefined ? 3 : o.maxRestarts;
  const T = load(targetDir);
  const opt = { cfl: T.s.opt.cfl, order: T.s.opt.order };
  const restarts = [];
  const t0 = process.hrtime.bigint();
  let log;
  for (;;) {
    log = { itTotal: 0, hist: [], grids: [] };
    if (attempt(coarseDir, T, every, cap, opt, make, load, log)) break;
    restarts.push({ at: log.itTotal, cfl: opt.cfl, order: opt.order });
    if (restarts.length > maxRestarts) {
      return { diverged: true, restarts, wall_s: Number(process.hrtime.bigint() - t0) / 1e9 };
    }
    opt.cfl = Math.max(0.3, opt.cfl - 0.2); opt.order = 1;   // src/app_cfd.js:407
  }
  const wall = Number(process.hrtime.bigint() - t0) / 1e9;
  const last = log.grids[log.grids.length - 1];
  return { grids: log.grids, iters: log.itTotal, F: last.F, res: last.res, wall_s: wall, hist: log.hist, restarts, opt };
}

if (require.main === module) {
  const [coarse, target, every, cap] = process.argv.slice(2);
  console.log(JSON.stringify(run(coarse, target, +(every || 250), +(cap || 60000))));
}
module.exports = { run };
