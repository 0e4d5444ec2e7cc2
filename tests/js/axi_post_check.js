'use strict';
/* Evaluates the JS host post-processing (3d_aero_lab_amd/js/cfd_post.js) on
 * the reference's own states from a golden case and prints the results.
 *   node tests/js/axi_post_check.js <golden dir> <tag>  */
const fs = require('fs'), path = require('path'), zlib = require('zlib');
const P = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'cfd_post.js'));
const { axiGrid } = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'cfd_axi.js'));
const [dir, tag] = process.argv.slice(2);
const m = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
const T = { f4: Float32Array, f8: Float64Array, u1: Uint8Array, i1: Int8Array };
const arr = n => { const a = m.arrays[n]; const b = zlib.gunzipSync(fs.readFileSync(path.join(dir, a.file))); const ab = b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength); return new T[a.dtype](ab); };
const s = m.scalars;
const grid = axiGrid(arr('xf'), arr('rf'));
const zones = {};
for (const k of ['solid', 'phi', 'ae', 'an', 'comp', 'wgt', 'lmix', 'tgt', 'rate', 'yrate', 'mfd', 'stream']) zones[k] = arr('z_' + k);
const comps = arr('comps');
const Q = [0, 1, 2, 3, 4].map(q => arr(`${tag}_Q${q}`));
const src = { S: arr(`${tag}_src`), F: arr(`${tag}_mfx`) };
const f = P.cfdForces(grid, zones, Q, comps, s.ctl.P0, s.hf, src);
const out = { F: f.F, Fb: f.Fb, Fw: f.Fw, mfuel: f.mfuel, pw: Array.from(f.pw) };
const sec = P.cfdSection(grid, zones, Q, s.ctl.comb.x1 - 0.004 * s.ctl.s, s.ctl.order.indexOf(s.ctl.comb.stream), src.F);
const f0 = P.cfdForces(grid, zones, Q, comps, s.ctl.P0, 0, src);
const c2 = Float64Array.from(comps);
const ctl = P.cfdControl(s.type || 'tj', s.ctl, grid, zones, Q, src.F, c2, f0.pw, s.ctl.etam, 250);
// display fields with and without the display solid rule, as sha256 per array
const crypto = require('crypto');
const sha = ta => crypto.createHash('sha256').update(Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength)).digest('hex');
zones.mat = arr('z_mat');
const derived = {};
for (const disp of [0, 1]) {
  const G = P.cfdDerived(grid, Object.assign({}, zones), Q, s.fl, null, !!disp);
  for (const k of ['reg', 'mat', ...P.DERIVED_KEYS]) derived[`der${disp}_${k}`] = sha(G[k]) === m.arrays[`der${disp}_${k}`].sha256;
}
console.log(JSON.stringify({ thrust: out, section: sec, control: { Tt4: ctl.Tt4, Tt7: ctl.Tt7, ratios: ctl.ratios, comps_after: Array.from(c2) }, derived }));
