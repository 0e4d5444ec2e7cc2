'use strict';
/* Argument guards of the Node-API addon. Prints one JSON line of
 * {check: error message or "no error"}.
 *   node tests/js/addon_guard_check.js cpu   foreign / wrong-kind handles (no device needed)
 *   node tests/js/addon_guard_check.js gpu   plus short arrays, use after dispose, double dispose */
const path = require('path');
const A = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'lib', 'aero3d.node'));
const out = {};
const expectThrow = (name, fn) => { try { fn(); out[name] = 'no error'; } catch (e) { out[name] = e.constructor.name + ': ' + e.message; } };

expectThrow('step_number_handle', () => A.step(123, 1));
expectThrow('step_object_handle', () => A.step({}, 1));
expectThrow('axi_step_number_handle', () => A.axiStep(7, 1));

if (process.argv[2] === 'gpu') {
  const K = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'cases.js'));
  const C = K.sphereCase();
  const g = C.g, geo = C.geo, N = g.N;
  const h = A.create(Float64Array.from(g.X.f), Float64Array.from(g.Y.f), Float64Array.from(g.Z.f), geo.t, geo.phi,
    geo.ap[0], geo.ap[1], geo.ap[2], geo.ap[3]);
  const f = n => new Float32Array(n);
  expectThrow('create_short_phi', () => A.create(Float64Array.from(g.X.f), Float64Array.from(g.Y.f), Float64Array.from(g.Z.f),
    geo.t, f(N - 1), geo.ap[0], geo.ap[1], geo.ap[2], geo.ap[3]));
  expectThrow('read_short_output', () => A.read(h, f(N), f(N), f(N - 1), f(N), f(N), null));
  expectThrow('write_short_input', () => A.write(h, f(N), f(N), f(N), f(N), f(10)));
  expectThrow('force_list_short_column', () => A.setForceList(h, new Int32Array([0, 1]), new Float64Array(2), new Float64Array(1),
    new Float64Array(2), new Float64Array(2)));
  expectThrow('axi_call_with_3d_handle', () => A.axiStep(h, 1));
  expectThrow('read_full_size', () => A.read(h, f(N), f(N), f(N), f(N), f(N), f(N)));
  A.dispose(h);
  expectThrow('step_after_dispose', () => A.step(h, 1));
  expectThrow('dispose_twice', () => A.dispose(h));
}
console.log(JSON.stringify(out));
