'use strict';
/* Command-line front end of the JS host layer.
 *
 *   node cli.js geom <case> <outdir>   build grid + cut-cell masks + force list,
 *                                      write raw little-endian arrays and meta.json
 *
 * <case>: sphere | transport:<cells>[:<delta>] | airliner|interceptor|fighter:<cells>[:<delta>]
 */
const fs = require('fs');
const path = require('path');
const A = require('./air_cfd.js');
const K = require('./cases.js');

function buildCase(spec) {
  const [kind, cells, delta] = spec.split(':');
  if (kind === 'sphere') return cells ? K.sphereCaseN(+cells) : K.sphereCase();
  if (kind === 'transport') { const C = K.transportCase(+cells, +(delta || 0)); C.fix = K.TRANSPORT_FIX; return C; }
  return K.aircraftCase(kind, +cells, +(delta || 0));
}

function writeGeom(C, dir) {
  fs.mkdirSync(dir, { recursive: true });
  const w = (n, ta) => fs.writeFileSync(path.join(dir, n + '.bin'), Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength));
  const g = C.g, geo = C.geo, fl = A.forceListArrays(A.cfdForceList(g, geo));
  w('fx', g.X.f); w('fy', g.Y.f); w('fz', g.Z.f);
  w('t', geo.t); w('phi', geo.phi); geo.ap.forEach((a, d) => w('ap' + d, a));
  w('fl_q', fl.q); w('fl_cx', fl.cx); w('fl_cz', fl.cz); w('fl_x', fl.x); w('fl_z', fl.z);
  const meta = { kind: C.kind, NX: g.NX, NY: g.NY, NZ: g.NZ, N: g.N, h: g.h, nSolid: geo.nSolid, nCut: geo.nCut,
    volW: geo.volW, volH: geo.volH, nForce: fl.n, ref: C.ref, xT: C.xT === undefined ? null : C.xT, fix: C.fix || null };
  fs.writeFileSync(path.join(dir, 'meta.json'), JSON.stringify(meta));
  return meta;
}

if (require.main === module && process.argv[2] === 'geom-many') {
  // node cli.js geom-many <outroot> <spec> ...: one directory per spec, <outroot>/<index>
  const root = process.argv[3], specs = process.argv.slice(4);
  specs.forEach((sp, k) => { const t0 = Date.now(); const m = writeGeom(buildCase(sp), path.join(root, String(k))); m.build_s = (Date.now() - t0) / 1000; m.spec = sp; fs.writeFileSync(path.join(root, String(k), 'meta.json'), JSON.stringify(m)); });
  console.log(JSON.stringify({ n: specs.length }));
} else if (require.main === module) {
  const [cmd, spec, out] = process.argv.slice(2);
  if (cmd !== 'geom' || !spec || !out) { console.error('usage: node cli.js geom <case> <outdir>'); process.exit(2); }
  const t0 = Date.now();
  const meta = writeGeom(buildCase(spec), out);
  meta.build_s = (Date.now() - t0) / 1000;
  fs.writeFileSync(path.join(out, 'meta.json'), JSON.stringify(meta));
  console.log(JSON.stringify(meta));
}
module.exports = { buildCase, writeGeom };
