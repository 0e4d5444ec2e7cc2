'use strict';
/* Builds aircraft cut-cell masks with the worker pool (3d_aero_lab_amd/js/geom_workers.js)
 * at several slab sizes and prints the sha256 of every mask array per build:
 *   node tests/js/geom_pool_check.js <workers> <type:cells:delta[:transport]> ...
 * -> [{"spec": ..., "slab": s, "t": sha, "phi": sha, "ap0".."ap3": sha, "nSolid", "nCut"}, ...] */
const path = require('path'), crypto = require('crypto');
const { GeomPool } = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'geom_workers.js'));
const B = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'bodies.js'));
const sha = ta => crypto.createHash('sha256').update(Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength)).digest('hex');

(async () => {
  const pool = new GeomPool(+process.argv[2]);
  const out = [];
  try {
    for (const spec of process.argv.slice(3)) {
      const [type, cells, delta, tr] = spec.split(':');
      for (const slab of [0, 1, 5]) {
        const C = await pool.aircraftCase(type, +cells, +delta, tr ? B.TRANSPORT_OVERRIDES : null, slab);
        const r = { spec, slab, t: sha(C.geo.t), phi: sha(C.geo.phi), nSolid: C.geo.nSolid, nCut: C.geo.nCut };
        C.geo.ap.forEach((a, d) => { r['ap' + d] = sha(a); });
        out.push(r);
      }
    }
  } finally {
    await pool.close();
  }
  console.log(JSON.stringify(out));
})().catch(e => { console.error(e); process.exit(1); });
