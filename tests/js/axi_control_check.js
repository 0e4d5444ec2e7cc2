'use strict';
/* The exit-temperature control of 3d_aero_lab_amd/js/cfd_post.js
 * (controlFromSections) fed the reference's combustor section and the
 * blade-row powers of its last thrust readback (tests/golden/axi_*); prints
 * the gains after the update, Tt4 and the spool ratios.
 *   node tests/js/axi_control_check.js <golden dir> */
const fs = require('fs'), path = require('path'), zlib = require('zlib');
const P = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'cfd_post.js'));
const dir = process.argv[2];
const m = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
const arr = n => { const a = m.arrays[n]; const b = zlib.gunzipSync(fs.readFileSync(path.join(dir, a.file))); return new Float64Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
const s = m.scalars;
const comps = arr('comps');
const pw = s.hist[s.hist.length - 1].pw;
const sec = s.section;
const ctl = P.controlFromSections(s.type || 'tj', s.ctl, st => (st.id === 4 ? { mdot: sec.mdot, Tt: sec.Tt } : null), comps, pw, s.ctl.etam, 250);
console.log(JSON.stringify({ comps_after: Array.from(comps), Tt4: ctl.Tt4, Tt7: ctl.Tt7, ratios: ctl.ratios }));
