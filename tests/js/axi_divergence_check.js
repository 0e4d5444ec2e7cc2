'use strict';
/* The divergence handling of the axisymmetric convergence loop (js/axi_seq.js,
 * src/app_cfd.js:377-378, 405-409) driven by a stand-in solver on the CPU:
 * the stand-in holds the case's initial state and, in the runs chosen, plants
 * a NaN density at cell 0 once it has stepped `at` iterations; its monitor
 * counts the cells k = 0, 7, 14, ... with density or energy not positive, as
 * the device monitor does, and reports a zero residual and no flow through
 * the control section. Prints JSON: the restarts, the options each solver was
 * created with, the outcome. */
const path = require('path');
const seq = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'axi_seq.js'));
const { loadCase } = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'axi_run.js'));

function standIn(plan) {
  const made = [];
  const make = (C, comps, Q0, opt) => {
    const k = made.length, N = C.grid.N;
    made.push(Object.assign({}, opt));
    const Q = Q0.map(a => Float32Array.from(a));
    let it = 0;
    return {
      step(n) { it += n; if (plan(k) !== null && it >= plan(k)) Q[0][0] = NaN; },
      read() { return Q.map(a => Float32Array.from(a)); },
      monitor() {
        let bad = 0;
        for (let c = 0; c < N; c += 7) if (!(Q[3][c] > 0) || !(Q[0][c] > 0)) bad++;
        return { F: 1, Fb: 1, Fw: 0, mfuel: 0, pw: new Float64Array(8), resSum2: 0, resN: 1, resValid: true, bad,
          sec: [{ mdot: 0, w: 0, wTt: 0 }, { mdot: 0, w: 0, wTt: 0 }] };
      },
      setComps() {}, monitorSections() {}, transferFrom() {}, dispose() {},
    };
  };
  return { make, made };
}

const dir = path.join(__dirname, '..', 'golden', 'axi_tj_xcoarse');
const out = {};
{ // the first solver diverges after 500 iterations; the restart runs clean
  const S = standIn(k => (k === 0 ? 500 : null));
  const r = seq.run('-', dir, 250, 60000, { makeSolver: S.make, loadCase: d => loadCase(d) });
  out.once = { restarts: r.restarts, made: S.made, iters: r.iters, diverged: !!r.diverged, opt: r.opt };
}
{ // every solver diverges: three restarts, then the run gives up
  const S = standIn(() => 250);
  const r = seq.run('-', dir, 250, 60000, { makeSolver: S.make, loadCase: d => loadCase(d), maxRestarts: 3 });
  out.always = { restarts: r.restarts, made: S.made, diverged: !!r.diverged };
}
console.log(JSON.stringify(out));
