'use strict';
/* Display fields and station sums of an axisymmetric state through the Node
 * addon (AxiHIP.derived, cfd_post.js stations over AxiHIP.sections), written
 * to files for the test to compare with the post-processing oracle.
 *   node tests/js/axi_display_dump.js <case dir> <state tag> <out dir>
 * out: der<disp>.f32 (13 planes), reg<disp>.u8, mat<disp>.u8, stations.json */
const fs = require('fs'), path = require('path');
const J = p => path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', p);
const P = require(J('cfd_post.js'));
const { AxiHIP } = require(J('cfd_axi.js'));
const { loadCase } = require(J('axi_run.js'));
const zlib = require('zlib');
const [dir, tag, out] = process.argv.slice(2);
const C = loadCase(dir), s = C.s;
const m = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
const arr = n => { const a = m.arrays[n]; const b = zlib.gunzipSync(fs.readFileSync(path.join(dir, a.file))); return new (a.dtype === 'f8' ? Float64Array : Uint8Array)(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
C.zones.mat = arr('z_mat');
const Q = [0, 1, 2, 3, 4].map(q => arr(`${tag}_Q${q}`));
const S = new AxiHIP(C.grid, C.zones, C.comps, s.fs, Q, Object.assign({}, s.opt, { hf: s.hf }));
for (const disp of [0, 1]) {
  const G = S.derived(s.fl, disp);
  fs.writeFileSync(path.join(out, `der${disp}.f32`), Buffer.from(G._f.buffer, G._f.byteOffset, G._f.byteLength));
  fs.writeFileSync(path.join(out, `reg${disp}.u8`), Buffer.from(G.reg.buffer));
  fs.writeFileSync(path.join(out, `mat${disp}.u8`), Buffer.from(G.mat.buffer));
}
// mass flux of a fresh solver is zero until it steps: one step, then the stations over the stepped state
S.step(1);
const st = P.stations(S, C.grid, C.zones, s.ctl.order, s.type || 'tj');
const Q1 = S.read(), mfx = S.mfx;
fs.writeFileSync(path.join(out, 'q1.f32'), Buffer.concat(Q1.map(a => Buffer.from(a.buffer, a.byteOffset, a.byteLength))));
fs.writeFileSync(path.join(out, 'mfx.f32'), Buffer.from(mfx.buffer, mfx.byteOffset, mfx.byteLength));
fs.writeFileSync(path.join(out, 'stations.json'), JSON.stringify(st));
S.dispose();
console.log(JSON.stringify({ ok: true, stations: st.length }));
