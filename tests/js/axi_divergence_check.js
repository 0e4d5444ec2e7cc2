'use strict';
/* The divergence handling of the axisymmetric convergence loop (js/axi_seq.js,
 * src/app_cfd.js:377-378, 405-409) driven by a stand-in solver on the CPU:
 * the stand-in holds the case's initial state (so the loop's own host
 * readback, forces and control run on real data) and, in the runs chosen,
 * plants a NaN density at cell 0 once it has stepped `at` iterations. Prints
 * JSON: the restarts, the options each solver was created with, the outcome. */
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
      src: new Float32Array(N), mfx: new Float32Array(N),
      step(n) { it += n; if (plan(k) !== null && it >= plan(k)) Q[0][0] = NaN; },
      read(out) { return Q.map((a, q) => { const o = out ? out[q] : new Float32Array(N); o.set(a); return o; }); },
      setComps() {}, monitorSections() {}, dispose() {},
    };
  };
  return { make, made };
}

const dir = path.join(__dirname, '..', 'golden', 'axi_tj_xcoarse');
const out = {};
{ // the first solver diverges after 500 iterations; the restart runs clean
  const S = standIn(k => (k === 0 ? 500 : null));
  const r = seq.run('-', dir, 250, 60000, true, { makeSolver: S.make, loadCase: d => loadCase(d) });
  out.once = { restarts: r.restarts, made: S.made, iters: r.iters, diverged: !!r.diverged, opt: r.opt };
}
{ // every solver diverges: three restarts, then the run gives up
  const S = standIn(() => 250);
  const r = seq.run('-', dir, 250, 60000, true, { makeSolver: S.make, loadCase: d => loadCase(d), maxRestarts: 3 });
  out.always = { restarts: r.restarts, made: S.made, diverged: !!r.diverged };
}
console.log(JSON.stringify(out));
