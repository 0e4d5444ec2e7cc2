'use strict';
/* The engine field model from Node (3d_aero_lab_amd/js/field_model.js).
 *   node tests/js/field_check.js tables   -> {"<model>": sha256 of fieldTable's bytes}   (no device)
 *   node tests/js/field_check.js device   -> JSON of the device results against the
 *        reference's own outputs: initialState on all twelve setups built from their
 *        engine descriptions (cfd_grid.js) at the reference's stride samples, the
 *        fixed states of CFD3.applyCond for the stock aircraft, FieldModel.sample at
 *        the seeded points through FieldHIP.sample / sampleMany.
 *   node tests/js/field_check.js fix <aircraft>   -> {F, fix} of fixedStates on the device */
const fs = require('fs'), path = require('path'), zlib = require('zlib'), crypto = require('crypto');
const ROOT = path.join(__dirname, '..', '..');
const FM = require(path.join(ROOT, '3d_aero_lab_amd', 'js', 'field_model.js'));
const GOLD = path.join(ROOT, 'tests', 'golden');
const manifest = d => JSON.parse(fs.readFileSync(path.join(GOLD, d, 'manifest.json'), 'utf8'));
const arr = (d, m, n) => { const a = m.arrays[n]; const b = zlib.gunzipSync(fs.readFileSync(path.join(GOLD, d, a.file))); return new Float64Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength)); };
const sha = ta => crypto.createHash('sha256').update(Buffer.from(ta.buffer, ta.byteOffset, ta.byteLength)).digest('hex');
const rel = (a, b) => (a === b || (a !== a && b !== b) ? 0 : Math.abs(a - b) / Math.max(Math.abs(b), 1e-300));

const F = manifest('axi_field'), AC = manifest('applycond');
const mode = process.argv[2];
const out = {};
if (mode === 'fix') {
  const c = AC.scalars.cases[process.argv[3]];
  const fm = new FM.FieldHIP(c.model);
  Object.assign(out, FM.fixedStates(fm, c.fs, c.M, c.alpha));
  fm.dispose();
} else if (mode === 'tables') {
  for (const [k, d] of Object.entries(F.scalars.models)) out[k] = sha(FM.fieldTable(d));
  for (const [k, c] of Object.entries(AC.scalars.cases)) out['applycond_' + k] = sha(FM.fieldTable(c.model));
} else {
  const G = require(path.join(ROOT, '3d_aero_lab_amd', 'js', 'cfd_grid.js'));
  const S = JSON.parse(fs.readFileSync(path.join(GOLD, 'axi_setups', 'manifest.json'), 'utf8')).setups;
  const models = {};
  const modelOf = t => (models[t] = models[t] || new FM.FieldHIP(F.scalars.models[t]));
  // initialState on every setup, from its own engine description
  out.initial = {};
  for (const [name, s] of Object.entries(S)) {
    const { grid, zones } = G.axiSetup(s.axispec, s.axispec.level);
    const Q = FM.initialState(grid, zones, modelOf(s.axispec.type), false);
    const ref = F.scalars.initial_state[name], smp = arr('axi_field', F, `q0s_${name}`), stride = ref.stride, ns = Math.ceil(grid.N / stride);
    let worst = 0, exact = 0;
    for (let q = 0; q < 5; q++) {
      if (sha(Q[q]) === ref.sha256[q]) exact++;
      for (let k = 0; k < ns; k++) worst = Math.max(worst, rel(Q[q][k * stride], smp[q * ns + k]));
    }
    out.initial[name] = { cells: grid.N, max_rel_sampled: worst, exact_planes: exact };
  }
  // CFD3.applyCond's freestream and fixed states
  out.applycond = {};
  for (const [k, c] of Object.entries(AC.scalars.cases)) {
    const fm = new FM.FieldHIP(c.model);
    const r = FM.fixedStates(fm, c.fs, c.M, c.alpha);
    let worst = 0;
    for (let q = 0; q < 5; q++) worst = Math.max(worst, rel(r.F[q], c.F[q]));
    for (let g = 0; g < 3; g++) for (let q = 0; q < 5; q++) worst = Math.max(worst, rel(r.fix[g][q], c.fix[g][q]));
    out.applycond[k] = { max_rel: worst };
    fm.dispose();
  }
  // FieldModel.sample at the seeded points: many at once, and the first 50 one by one
  out.sample = {};
  for (const name of Object.keys(F.scalars.models)) {
    const fm = new FM.FieldHIP(F.scalars.models[name]);
    const x = arr('axi_field', F, `${name}_x`), r = arr('axi_field', F, `${name}_r`), th = arr('axi_field', F, `${name}_th`), ref = arr('axi_field', F, `${name}_out`);
    const v = fm.sampleMany(x, r, th);
    let worst = 0, regions = 0;
    for (let p = 0; p < x.length; p++) {
      if (v[14 * p] !== ref[14 * p]) regions++;
      for (let q = 1; q < 14; q++) worst = Math.max(worst, rel(v[14 * p + q], ref[14 * p + q]));
    }
    let single = 0;
    for (let p = 0; p < 50; p++) {
      const o = fm.sample(x[p], r[p], th[p]);
      if ((o.solid ? 1 : o.region === 'ext' ? 0 : 2 + fm.order.indexOf(o.stream)) !== ref[14 * p]) single++;
      if (!o.solid) FM.SAMPLE_KEYS.forEach((k, q) => { if (o[k] !== v[14 * p + 1 + q]) single++; });
    }
    out.sample[name] = { n: x.length, region_mismatch: regions, max_rel: worst, single_vs_many_mismatch: single };
    fm.dispose();
  }
  for (const m of Object.values(models)) m.dispose();
}
console.log(JSON.stringify(out));
