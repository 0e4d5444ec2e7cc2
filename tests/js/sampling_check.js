'use strict';
/* Display sampling of 3d_aero_lab_amd/js/air_cfd.js (cfdStencil, cfdSample,
 * cfdState, cfdVelAt) on the reference's fighter state and probe points of
 * tests/golden/sampling; prints which outputs equal the reference's bit for bit.
 *   node tests/js/sampling_check.js <golden dir> */
const fs = require('fs'), path = require('path'), zlib = require('zlib');
const A = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'air_cfd.js'));
const dir = process.argv[2];
const m = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
const T = { f4: Float32Array, f8: Float64Array, u1: Uint8Array, i4: Int32Array };
const arr = n => { const a = m.arrays[n]; const b = zlib.gunzipSync(fs.readFileSync(path.join(dir, a.file))); return new T[a.dtype](b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
const s = m.scalars;
const g = A.gridFromFaces(arr('fx'), arr('fy'), arr('fz'));
const t = arr('t'), H = { r: arr('it200_r'), ru: arr('it200_ru'), rv: arr('it200_rv'), rw: arr('it200_rw'), E: arr('it200_E') };
const pts = arr('pts'), n = pts.length / 3;
const idx = new Int32Array(8 * n), w = new Float32Array(8 * n), ok = new Uint8Array(n), got = new Uint8Array(n);
const Wv = new Float64Array(5 * n), st = new Float64Array(12 * n), vel = new Float64Array(3 * n), vok = new Uint8Array(n);
const W = [0, 0, 0, 0, 0], o = {}, vo = [0, 0, 0];
for (let p = 0; p < n; p++) {
  const x = pts[3 * p], y = pts[3 * p + 1], z = pts[3 * p + 2];
  ok[p] = A.cfdStencil(g, t, x, y, z, idx, w, 8 * p) ? 1 : 0;
  if (A.cfdSample(H, idx, w, 8 * p, y < 0 ? -1 : 1, W)) {
    got[p] = 1; for (let c = 0; c < 5; c++) Wv[5 * p + c] = W[c];
    A.cfdState(s.fs, W, o); s.keys.forEach((k, c) => { st[12 * p + c] = o[k]; });
  }
  if (A.cfdVelAt(g, t, H, s.M, x, y, z, vo)) { vok[p] = 1; vel[3 * p] = vo[0]; vel[3 * p + 1] = vo[1]; vel[3 * p + 2] = vo[2]; }
}
const same = (a, b) => Buffer.from(a.buffer, a.byteOffset, a.byteLength).equals(Buffer.from(b.buffer, b.byteOffset, b.byteLength));
console.log(JSON.stringify({ n, idx: same(idx, arr('idx')), w: same(w, arr('w')), ok: same(ok, arr('ok')), got: same(got, arr('got')),
  W: same(Wv, arr('W')), state: same(st, arr('state')), vel: same(vel, arr('vel')), vok: same(vok, arr('vok')) }));
