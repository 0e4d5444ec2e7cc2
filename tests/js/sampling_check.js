'use strict';
/* Display sampling on the device through the Node addon (CfdHIP.probe of
 * 3d_aero_lab_amd/js/hip_solver.js) on the reference's fighter state and the
 * probe points of tests/golden/sampling; prints which outputs equal the
 * reference's (cfdStencil / cfdSample / cfdState / cfdVelAt) bit for bit and
 * the largest relative difference of the fields that go through sqrt / pow.
 *   node tests/js/sampling_check.js <sampling dir> <fighter30k dir> */
const fs = require('fs'), path = require('path'), zlib = require('zlib');
const A = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'air_cfd.js'));
const { CfdHIP, PROBE_KEYS } = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'hip_solver.js'));
const T = { f4: Float32Array, f8: Float64Array, u1: Uint8Array, i4: Int32Array };
function loader(dir) {
  const m = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
  const arr = n => { const a = m.arrays[n]; const b = zlib.gunzipSync(fs.readFileSync(path.join(dir, a.file))); return new T[a.dtype](b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
  return { m, s: m.scalars, arr };
}
const S = loader(process.argv[2]), F = loader(process.argv[3]);
const g = A.gridFromFaces(F.arr('fx'), F.arr('fy'), F.arr('fz'));
const geo = { t: F.arr('t'), phi: F.arr('phi'), ap: [0, 1, 2, 3].map(d => F.arr('ap' + d)) };
const s = new CfdHIP(g, geo);
const H = { r: S.arr('it200_r'), ru: S.arr('it200_ru'), rv: S.arr('it200_rv'), rw: S.arr('it200_rw'), E: S.arr('it200_E') };
// write() re-applies the fixed states (cell types 3, 4, 5): give it the run's own, as the state holds them
const fixOf = ty => { const k = Math.max(0, geo.t.indexOf(ty)); return [H.r[k], H.ru[k], H.rv[k], H.rw[k], H.E[k]]; };
s.setFix([fixOf(3), fixOf(4), fixOf(5)]);
s.write(H);
const pts = S.arr('pts'), n = pts.length / 3;
const o = s.probe(pts, S.s.fs, S.s.M);
s.dispose();
const bytes = a => Buffer.from(a.buffer, a.byteOffset, a.byteLength);
const same = (a, b) => bytes(a).equals(bytes(b));
const ok = S.arr('ok'), vok = S.arr('vok'), W = S.arr('W'), st = S.arr('state'), vel = S.arr('vel');
let Wsame = true, velSame = true, stExact = true;
const worst = {};
const exact = new Set(['P', 'T', 'rho', 'Cp', 'Vx', 'Vy', 'Vz']);
for (let p = 0; p < n; p++) {
  if (ok[p]) {
    for (let c = 0; c < 5; c++) if (!Object.is(o.W[5 * p + c], W[5 * p + c])) Wsame = false;
    PROBE_KEYS.forEach((k, c) => {
      const a = o.state[12 * p + c], b = st[12 * p + c];
      if (exact.has(k)) { if (!Object.is(a, b)) stExact = false; }
      else worst[k] = Math.max(worst[k] || 0, Math.abs(a - b) / Math.abs(b));
    });
  }
  if (vok[p]) for (let c = 0; c < 3; c++) if (!Object.is(o.vel[3 * p + c], vel[3 * p + c])) velSame = false;
}
console.log(JSON.stringify({ n, idx: same(o.idx, S.arr('idx')), w: same(o.w, S.arr('w')), ok: same(o.ok, ok), got: same(o.ok, S.arr('got')),
  W: Wsame, stateExact: stExact, vel: velSame, vok: same(o.vok, vok), keys: PROBE_KEYS.join() === S.s.keys.join(), worst }));
