'use strict';
/* Checkpoint files of 3d_aero_lab_amd/js/cfd_axi.js against the reference's
 * (tests/golden/axi_ckpt: its save() on the TJ xcoarse state after 10
 * iterations): same bytes, and loading gives the planes back.
 *   node tests/js/axi_ckpt_check.js <golden root> <scratch dir> */
const fs = require('fs'), path = require('path'), zlib = require('zlib'), crypto = require('crypto');
const X = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'cfd_axi.js'));
const [root, tmp] = process.argv.slice(2);
const ref = JSON.parse(fs.readFileSync(path.join(root, 'axi_ckpt', 'manifest.json'), 'utf8'));
const dir = path.join(root, 'axi_tj_xcoarse'), m = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
const arr = n => { const a = m.arrays[n]; const b = zlib.gunzipSync(fs.readFileSync(path.join(dir, a.file))); return new Float64Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
const Q = [0, 1, 2, 3, 4].map(q => arr(`it10_Q${q}`)), comps = arr('comps');
const f = path.join(tmp, 'ckpt.bin');
X.saveCheckpoint(f, Q, comps);
const bin = fs.readFileSync(f), js = fs.readFileSync(f + '.comps.json', 'utf8');
const L = X.loadCheckpoint(f, Q[0].length, 32), bad = X.loadCheckpoint(f, Q[0].length + 1, 32);
console.log(JSON.stringify({ bytes: bin.length === ref.bin_bytes, sha: crypto.createHash('sha256').update(bin).digest('hex') === ref.bin_sha256,
  comps: js === ref.comps_json, roundtrip: L.Q.every((a, q) => Buffer.from(a.buffer).equals(Buffer.from(Q[q].buffer))) && L.comps.every((v, k) => v === comps[k]),
  wrong_size_ignored: bad.Q === null && bad.comps !== null }));
