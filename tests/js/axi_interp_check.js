'use strict';
/* interpState of 3d_aero_lab_amd/js/cfd_grid.js on the reference's xcoarse
 * states onto the medium grids (tests/golden/axi_interp): sha256 of each plane.
 *   node tests/js/axi_interp_check.js <golden root>   -> {"tj": [true x5], ...} */
const fs = require('fs'), path = require('path'), zlib = require('zlib'), crypto = require('crypto');
const G = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'cfd_grid.js'));
const { axiGrid } = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'cfd_axi.js'));
const root = process.argv[2];
const T = { f4: Float32Array, f8: Float64Array, u1: Uint8Array, i1: Int8Array };
function load(name) {
  const dir = path.join(root, name), m = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
  const arr = n => { const a = m.arrays[n]; const b = zlib.gunzipSync(fs.readFileSync(path.join(dir, a.file))); return new T[a.dtype](b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
  const zones = { solid: arr('z_solid') };
  return { m, arr, grid: axiGrid(arr('xf'), arr('rf')), zones };
}
const sha = ta => crypto.createHash('sha256').update(Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength)).digest('hex');
const ref = JSON.parse(fs.readFileSync(path.join(root, 'axi_interp', 'manifest.json'), 'utf8')).interp;
const out = {};
for (const [type, r] of Object.entries(ref)) {
  const A = load(`axi_${type}_${r.from}`), B = load(`axi_${type}_${r.to}`);
  const QA = [0, 1, 2, 3, 4].map(q => A.arr(`it${r.iters}_Q${q}`));
  const Q = G.interpState(A.grid, A.zones, QA, B.grid, B.zones, B.m.scalars.fl);
  out[type] = Q.map((a, q) => sha(a) === r.sha256[q]);
}
console.log(JSON.stringify(out));
