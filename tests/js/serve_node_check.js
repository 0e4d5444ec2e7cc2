'use strict';
/* The served and batched forms driven from Node (js/hip_solver.js Server /
 * Batch over the addon's serverTickAsync / batchStepAsync): requests of three
 * bodies under both convergence rules, submitted in two waves, each result
 * compared with a plain CfdHIP solve read back by CfdMonitor under the same
 * rule (iterations, residual, coefficients: the same bits); stepAsync and
 * batchStepAsync against their synchronous forms. Prints one JSON line. */
const path = require('path');
const J = path.join(__dirname, '..', '..', '3d_aero_lab_amd', 'js');
const { CfdHIP, CfdMonitor, ConvRule, RULE_RES_DROP, RULE_REFERENCE, Batch, Server } = require(path.join(J, 'hip_solver.js'));
const K = require(path.join(J, 'cases.js'));
const A = require(path.join(J, 'air_cfd.js'));

function freestream(M, alphaDeg) {
  const a = alphaDeg * Math.PI / 180;
  return [1, M * Math.cos(a), 0, M * Math.sin(a), 1 / (A.CG * (A.CG - 1)) + 0.5 * M * M];
}

(async () => {
  const cases = [K.sphereCaseN(30000), K.aircraftCase('interceptor', 30000, -2), K.transportCase(60000)];
  cases[2].fix = K.TRANSPORT_FIX;
  const lists = cases.map(C => A.cfdForceList(C.g, C.geo));
  const reqs = [];
  cases.forEach((C, b) => {
    reqs.push({ b, M: 0.45 + 0.1 * b, alpha: 1 + b, rule: RULE_RES_DROP, resDrop: 2e-2, maxIters: 430 });
    reqs.push({ b, M: 0.6, alpha: 2.5 - b, rule: RULE_REFERENCE, resDrop: 0, maxIters: 700 });
  });
  const srv = new Server(4 * 2 ** 30, 4, 50);
  const gids = cases.map((C, b) => srv.addGeometry(C.g, C.geo, lists[b]));
  const submit = (k) => {
    const r = reqs[k], C = cases[r.b], F = freestream(r.M, r.alpha);
    srv.submit({ geometry: gids[r.b], F, fix: C.fix || [F, F, F], cfl: 0.5, order: 2, alpha: r.alpha, M: r.M, S: C.ref.S, xcg: C.ref.xcg,
      zcg: 0, cref: C.ref.c, rule: r.rule, resDrop: r.resDrop, maxIters: r.maxIters, arrival: srv.clock(), tag: k });
  };
  for (let k = 0; k < 3; k++) submit(k);
  const done = {};
  let ticks = 0;
  while (Object.keys(done).length < reqs.length) {
    if (ticks === 4) for (let k = 3; k < reqs.length; k++) submit(k);
    await srv.tickAsync();
    ticks++;
    for (const o of srv.poll()) done[o.tag] = o;
    if (ticks > 5000) throw new Error('served requests did not finish');
  }
  const stats = srv.stats();
  srv.dispose();
  const cmp = [];
  let ok = true;
  reqs.forEach((r, k) => {
    const C = cases[r.b], F = freestream(r.M, r.alpha);
    const s = new CfdHIP(C.g, C.geo);
    s.opt.cfl = 0.5; s.setFree(F); s.setFix(C.fix || [F, F, F]); s.reset(); s.setForceList(lists[r.b]);
    const m = new CfdMonitor(s, { alpha: r.alpha, M: r.M, S: C.ref.S, xcg: C.ref.xcg, zcg: 0, cref: C.ref.c }, undefined,
      new ConvRule(r.rule, r.resDrop, r.maxIters));
    const last = m.solve(50);
    s.dispose();
    const d = done[k];
    const same = d.iters === last.it && d.CL === last.CL && d.CD === last.CD && d.Cm === last.Cm && d.res === last.res &&
      !!d.converged === m.rule.conv && d.nreads === m.hist.length;
    ok = ok && same;
    cmp.push({ k, body: r.b, rule: r.rule, served: [d.iters, d.converged, d.CL], plain: [last.it, m.rule.conv, last.CL], same });
  });
  // stepAsync against step, batchStepAsync against batchStep
  const C = cases[0], F = freestream(0.5, 0);
  const a = new CfdHIP(C.g, C.geo), b = new CfdHIP(C.g, C.geo);
  for (const s of [a, b]) { s.opt.cfl = 0.5; s.setFree(F); s.setFix([F, F, F]); s.reset(); }
  a.step(37);
  await b.stepAsync(37);
  const Ha = Float32Array.from(a.read().E), Hb = Float32Array.from(b.read().E);
  let busyThrows = false;
  const p = b.stepAsync(5);
  try { b.read(); } catch (e) { busyThrows = /busy/.test(e.message); }
  await p;
  a.dispose(); b.dispose();
  const stepSame = Buffer.compare(Buffer.from(Ha.buffer), Buffer.from(Hb.buffer)) === 0;
  const B1 = new Batch(2 ** 30, 4), B2 = new Batch(2 ** 30, 4);
  const g1 = B1.addGeometry(C.g, C.geo, lists[0]), g2 = B2.addGeometry(C.g, C.geo, lists[0]);
  const s1 = B1.add(g1, F, [F, F, F], 0.5), s2 = B2.add(g2, F, [F, F, F], 0.5);
  B1.step(20);
  await B2.stepAsync(20);
  const r1 = B1.read(s1).r, r2 = B2.read(s2).r;
  const batchSame = Buffer.compare(Buffer.from(r1.buffer), Buffer.from(r2.buffer)) === 0 && B2.iter(s2) === 20;
  B1.dispose(); B2.dispose();
  console.log(JSON.stringify({ ok: ok && stepSame && batchSame && busyThrows, served: cmp, ticks, stats, stepSame, batchSame, busyThrows }));
})().catch(e => { console.error(e); process.exit(1); });
