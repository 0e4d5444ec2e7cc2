'use strict';
/* The reference's CPU driver flow (tools/test_cfd3.js:17-24) on the HIP solver:
 * aircraft -> cfdGrid -> cfdGeom -> cfdForceList -> new CfdHIP -> setFree /
 * setFix([F,F,F]) / reset -> step(50) + force coefficients (device FP64
 * reduction, the reference's cfdForces on the device), printed as JSON.
 *   node tests/js/cfd3_hip_run.js <type> <cells> <iters> <M> <alpha> <delta> <cfl>
 */
const path = require('path');
const PKG = path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js');
const A = require(path.join(PKG, 'air_cfd.js'));
const K = require(path.join(PKG, 'cases.js'));
const { CfdHIP } = require(path.join(PKG, 'hip_solver.js'));

const [type, cells, iters, M, alpha, delta, cfl] = process.argv.slice(2).map((v, i) => (i === 0 ? v : +v));
const C = K.aircraftCase(type, cells, delta);
const g = C.g, geo = C.geo, list = A.cfdForceList(g, geo);
const s = new CfdHIP(g, geo);
s.opt.cfl = cfl;
const F = A.freestream(M, alpha);
s.setFree(F); s.setFix([F, F, F]); s.reset();
s.setForceList(list);
const out = [];
for (let it = 50; it <= iters; it += 50) {
  s.step(50);
  out.push({ it, dev: s.forces(alpha, M, C.ref.S, C.ref.xcg, 0, C.ref.c) });
}
const cdi = s.trefftz(C.xT, alpha, M, C.ref.S);
s.dispose();
console.log(JSON.stringify({ N: g.N, hist: out, cdi }));
