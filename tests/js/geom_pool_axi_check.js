'use strict';
/* Axisymmetric setups built on the worker pool (GeomPool.axiSetup) at several
 * column-slab sizes, compared with the reference's arrays (sha256 of the bytes).
 *   node tests/js/geom_pool_axi_check.js <workers> <golden dir with manifest of setups>
 * -> {"<setup>/<slab>": true | [names of arrays that differ], ...} */
const fs = require('fs'), path = require('path'), crypto = require('crypto');
const { GeomPool } = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'geom_workers.js'));
const ZONE_KEYS = ['solid', 'phi', 'ae', 'an', 'comp', 'wgt', 'lmix', 'tgt', 'rate', 'yrate', 'mfd', 'stream', 'mat'];
const sha = ta => crypto.createHash('sha256').update(Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength)).digest('hex');

(async () => {
  const pool = new GeomPool(+process.argv[2]);
  const m = JSON.parse(fs.readFileSync(path.join(process.argv[3], 'manifest.json'), 'utf8'));
  const res = {};
  try {
    for (const [name, S] of Object.entries(m.setups)) {
      for (const slab of [0, 1, 7]) {
        const { grid, zones, comps } = await pool.axiSetup(S.axispec, S.axispec.level, slab);
        const bad = [];
        if (grid.W !== S.W || grid.H !== S.H) bad.push('size');
        if (sha(grid.xf) !== S.sha256.xf) bad.push('xf');
        if (sha(grid.rf) !== S.sha256.rf) bad.push('rf');
        if (sha(comps) !== S.sha256.comps) bad.push('comps');
        for (const k of ZONE_KEYS) if (sha(zones[k]) !== S.sha256['z_' + k]) bad.push(k);
        res[name + '/' + slab] = bad.length ? bad : true;
      }
    }
  } finally {
    await pool.close();
  }
  console.log(JSON.stringify(res));
})().catch(e => { console.error(e); process.exit(1); });
