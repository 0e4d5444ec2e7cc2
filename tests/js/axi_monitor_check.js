'use strict';
/* Device readback monitor (AxiHIP.monitor) against the host functions on the
 * same state: cfdForces with the solver's sources, the density residual of the
 * readback and cfdSection at the control stations.
 *   node tests/js/axi_monitor_check.js <case dir> <iters>
 * -> {"forces": max rel diff, "residual": rel diff, "sections_exact": bool, ...} */
const path = require('path');
const J = p => path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js', p);
const P = require(J('cfd_post.js'));
const { AxiHIP } = require(J('cfd_axi.js'));
const { loadCase } = require(J('axi_run.js'));
const { residual } = require(J('axi_seq.js'));

const [dir, iters] = process.argv.slice(2);
const C = loadCase(dir), s = C.s;
const comps = Float64Array.from(C.comps);
const S = new AxiHIP(C.grid, C.zones, comps, s.fs, C.Q0, Object.assign({}, s.opt, { hf: s.hf }));
const st = P.controlStations(Object.assign({}, s.ctl, { abOn: true, ab: s.ctl.ab || null }));
S.monitorSections(st.map(x => P.sectionRows(C.grid, C.zones, x.x, x.si)));
const rel = (a, b) => Math.abs(a - b) / Math.max(Math.abs(b), 1e-30);
S.step(+iters);
const m1 = S.monitor(comps, s.ctl.P0);
const Q1 = S.read();
const rho1 = Float32Array.from(Q1[0]);
S.step(50);
const m2 = S.monitor(comps, s.ctl.P0);
const Q2 = S.read();
const f = P.cfdForces(C.grid, C.zones, Q2, comps, s.ctl.P0, s.hf, { S: S.src, F: S.mfx });
const out = { first_res_valid: m1.resValid, second_res_valid: m2.resValid, nsec: m2.sec.length };
let fr = 0;
for (const k of ['F', 'Fb', 'Fw', 'mfuel']) fr = Math.max(fr, rel(m2[k], f[k]));
for (let q = 0; q < 8; q++) if (f.pw[q] !== 0 || m2.pw[q] !== 0) fr = Math.max(fr, rel(m2.pw[q], f.pw[q]));
out.forces = fr;
out.F = f.F;
const rh = residual(C.zones, Q2[0], rho1, 50), rd = Math.sqrt(m2.resSum2 / Math.max(1, m2.resN)) / 50;
out.residual = rel(rd, rh);
out.res = rh;
out.sections_exact = st.every((x, q) => {
  const h = P.cfdSection(C.grid, C.zones, Q2, x.x, x.si, S.mfx), d = P.sectionFromSums(m2.sec[q]);
  return h.mdot === d.mdot && h.Tt === d.Tt;
});
out.sections = st.map((x, q) => ({ id: x.id, host: P.cfdSection(C.grid, C.zones, Q2, x.x, x.si, S.mfx).Tt, dev: P.sectionFromSums(m2.sec[q]).Tt }));
S.dispose();
console.log(JSON.stringify(out));
