'use strict';
/* The JS convergence rule (js/hip_solver.js ConvRule) over the reference
 * controller's readback histories (tests/golden/readback_*): prints the conv
 * flag after every readback per case, and edge-case results, as JSON. */
const fs = require('fs');
const path = require('path');
const { ConvRule, RULE_REFERENCE, RULE_RES_DROP } = require(path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', 'hip_solver.js'));

const out = {};
for (const c of ['readback_sphere', 'readback_fighter30k', 'readback_transport60k']) {
  const m = JSON.parse(fs.readFileSync(path.join(__dirname, '..', 'golden', c, 'manifest.json'), 'utf8'));
  const R = new ConvRule(RULE_REFERENCE, 0, 30000);
  out[c] = m.scalars.reads.map(x => { R.update(x.res === null ? NaN : x.res, x.CL, x.it); return R.conv; });
}
const edge = {};
{
  const R = new ConvRule(RULE_RES_DROP, 1e-4, 30000);
  [NaN, 1e-3, 5e-4, NaN].forEach((r, k) => R.update(r, 0.1, 50 * k));
  edge.nan_res_diverged = R.diverged && R.done && !R.conv;
}
{
  const R = new ConvRule(RULE_REFERENCE, 0, 30000);
  const res = [NaN]; for (let k = 0; k < 11; k++) res.push(1e-3 * 0.25 ** k);
  const flags = res.map((r, k) => { R.update(r, k === 7 ? NaN : 0.3, 50 * k); return R.conv; });
  edge.nan_cl_no_conv = JSON.stringify(flags) === JSON.stringify([false, false, false, false, false, false, true, false, false, false, false, true]);
}
{
  const R = new ConvRule(RULE_REFERENCE, 0, 100);
  const d = [NaN, 1e-3, 1e-3, 1e-3].map((r, k) => R.update(r, 0.1, 50 * k));
  const Q = new ConvRule(RULE_RES_DROP, 1e-4, 100);
  const e = [NaN, 1e-3, 1e-3].map((r, k) => Q.update(r, 0.1, 50 * k));
  edge.cap = JSON.stringify(d) === '[false,false,false,true]' && JSON.stringify(e) === '[false,false,true]';
}
out.edge = edge;
console.log(JSON.stringify(out));
