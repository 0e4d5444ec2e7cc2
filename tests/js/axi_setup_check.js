'use strict';
/* Builds axisymmetric setups (grid faces, zones, component constants) from the
 * engine descriptions of a golden case with 3d_aero_lab_amd/js/cfd_grid.js and
 * compares every array with the reference's (sha256 of the bytes).
 *   node tests/js/axi_setup_check.js <golden dir>   -> {"<setup>": {"xf": true, ...}, ...}
 * The directory holds one setup (axi_<type>_<level>) or many (axi_setups). */
const fs = require('fs'), path = require('path'), crypto = require('crypto');
const G = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'cfd_grid.js'));
const ZONE_KEYS = ['solid', 'phi', 'ae', 'an', 'comp', 'wgt', 'lmix', 'tgt', 'rate', 'yrate', 'mfd', 'stream', 'mat'];
const sha = ta => crypto.createHash('sha256').update(Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength)).digest('hex');

function check(spec, W, H, ref) {
  const { grid, zones, comps } = G.axiSetup(spec, spec.level);
  const out = { W: grid.W === W, H: grid.H === H, xf: sha(grid.xf) === ref.xf, rf: sha(grid.rf) === ref.rf, comps: sha(comps) === ref.comps };
  for (const k of ZONE_KEYS) out['z_' + k] = sha(zones[k]) === ref['z_' + k];
  return out;
}

const dir = process.argv[2];
const m = JSON.parse(fs.readFileSync(path.join(dir, 'manifest.json'), 'utf8'));
const res = {};
if (m.setups) {
  for (const [name, S] of Object.entries(m.setups)) res[name] = check(S.axispec, S.W, S.H, S.sha256);
} else {
  const ref = {};
  for (const k of ['xf', 'rf', 'comps', ...ZONE_KEYS.map(k => 'z_' + k)]) ref[k] = m.arrays[k].sha256;
  res[m.case] = check(m.scalars.axispec, m.scalars.W, m.scalars.H, ref);
}
console.log(JSON.stringify(res));
