'use strict';
/* ============================================================================
 * serve_stream.js — the served form on a seeded stream of solve requests,
 * driven from Node with the cut-cell build in the request path.
 *
 *   node serve_stream.js [n=10000] [rate=20] [seed=20261204] [maxIters=2000] [resDrop=1e-4] [workers=7]
 *
 * Requests arrive as a Poisson stream (mean `rate` per second). Each is a body
 * (sphere, the 40 m transport, or the swept-delta single-fuselage
 * interceptor, 1/3 each) at a grid level of 30k / 120k / 500k cells (weights
 * 0.5 / 0.3 / 0.2) with M ~ U[0.3, 0.85] and alpha ~ U[-2, 8] deg, from one
 * mulberry32 seed (the sequence bench.py's serve_workload draws). A request
 * whose body and level have no geometry on the device yet waits for its
 * cut-cell build (aircraft on a GeomPool of worker threads, the sphere on the
 * calling thread), which its latency includes; the geometry then stays
 * registered for later requests. The server (Server, over aero3d_server_*)
 * admits by device-memory footprint, steps every resident solve, reads each
 * back every 50 iterations and ends it at res / res0 <= resDrop or maxIters.
 * Ticks run asynchronously (serverTickAsync); the loop submits arrivals
 * between ticks. Prints a progress line every 10 s and one JSON line at the
 * end: latency percentiles (arrival to result, build included), solves/s,
 * device-busy fraction, cell-updates/s, build times.
 * ========================================================================== */
const path = require('path');
const A = require('./air_cfd.js');
const B = require('./bodies.js');
const K = require('./cases.js');
const { GeomPool } = require('./geom_workers.js');
const { Server, RULE_RES_DROP } = require('./hip_solver.js');

const LEVELS = [30000, 120000, 500000];
const BODIES = ['sphere', 'transport', 'interceptor'];

function stream(n, rate, seed) {
  const rnd = B.mulberry32(seed);
  const out = [];
  let t = 0;
  for (let k = 0; k < n; k++) {
    t += -Math.log(1 - rnd()) / rate;
    const b = Math.floor(rnd() * 3), u = rnd(), lv = u < 0.5 ? 0 : u < 0.8 ? 1 : 2;
    const M = 0.3 + 0.55 * rnd(), alpha = -2 + 10 * rnd();
    out.push({ tag: k, body: BODIES[b], cells: LEVELS[lv], M, alpha, arrival: t });
  }
  return out;
}

function freestream(M, alphaDeg) {
  const a = alphaDeg * Math.PI / 180;
  return [1, M * Math.cos(a), 0, M * Math.sin(a), 1 / (A.CG * (A.CG - 1)) + 0.5 * M * M];
}

async function main() {
  const [n = 10000, rate = 20, seed = 20261204, maxIters = 2000, resDrop = 1e-4, workers = 7] = process.argv.slice(2).map(Number);
  const reqs = stream(n, rate, seed);
  const pool = new GeomPool(workers);
  const srv = new Server(64 * 2 ** 30, 512, 50);
  const geoms = new Map();          // key -> {gid, ready: Promise, built: {ref, fix}, build_s}
  const builds = [];
  let buildQueue = Promise.resolve();
  function geometry(body, cells) {
    const key = body + ':' + cells;
    if (geoms.has(key)) return geoms.get(key);
    const e = { gid: -1, info: null };
    // builds run one after another (the pool takes one build at a time)
    e.ready = buildQueue = buildQueue.then(async () => {
      const t0 = process.hrtime.bigint();
      let C;
      if (body === 'sphere') C = K.sphereCaseN(cells);
      else if (body === 'transport') { C = await pool.transportCase(cells); C.fix = K.TRANSPORT_FIX; }
      else C = await pool.aircraftCase(body, cells);
      e.list = A.cfdForceList(C.g, C.geo);
      e.C = C;                       // registered on the device between two ticks (the server handle is busy during one)
      e.info = { ref: C.ref, fix: C.fix || null, N: C.g.N };
      const s = Number(process.hrtime.bigint() - t0) / 1e9;
      builds.push({ key, cells: C.g.N, build_s: s });
    });
    geoms.set(key, e);
    return e;
  }
  const t0 = srv.clock();
  const now = () => srv.clock() - t0;
  const waiting = [];               // arrived, geometry not ready yet
  const lat = [];
  let next = 0, done = 0, lastPrint = 0, converged = 0, failed = 0, iters = 0;
  const submit = (r, e) => {
    const F = freestream(r.M, r.alpha), ref = e.info.ref;
    srv.submit({ geometry: e.gid, F, fix: e.info.fix || [F, F, F], cfl: 0.5, order: 2, alpha: r.alpha, M: r.M, S: ref.S, xcg: ref.xcg,
      zcg: 0, cref: ref.c, rule: RULE_RES_DROP, resDrop, maxIters, arrival: t0 + r.arrival, tag: r.tag });
  };
  while (done < n) {
    const t = now();
    while (next < n && reqs[next].arrival <= t) {
      const r = reqs[next++];
      const e = geometry(r.body, r.cells);
      waiting.push({ r, e });
    }
    for (const e of geoms.values()) if (e.C && e.gid < 0) { e.gid = srv.addGeometry(e.C.g, e.C.geo, e.list); e.C = null; e.list = null; }
    for (let k = 0; k < waiting.length;) {
      if (waiting[k].e.gid >= 0) { submit(waiting[k].r, waiting[k].e); waiting.splice(k, 1); } else k++;
    }
    const st = await srv.tickAsync();
    for (const o of srv.poll()) {
      done++; lat.push(o.tDone - o.arrival); iters += o.iters;
      if (o.converged) converged++;
      if (o.failed) failed++;
    }
    if (st.active === 0 && st.pending === 0) {
      // idle: wait for the next arrival or a pending build
      const wait = waiting.length ? 2 : next < n ? Math.max(0, (reqs[next].arrival - now()) * 1000) : 1;
      await new Promise(res => setTimeout(res, Math.min(wait, 50)));
    }
    if (t - lastPrint > 10) {
      lastPrint = t;
      const s = srv.stats();
      console.log(`[serve_stream] t ${t.toFixed(1)} s: arrived ${next}, done ${done}, resident ${st.active}, queued ${st.pending}, ` +
        `waiting for a build ${waiting.length}, device busy ${(s.gpuS / Math.max(t, 1e-9)).toFixed(3)}`);
    }
  }
  const wall = now();
  const s = srv.stats();
  lat.sort((a, b) => a - b);
  const pct = p => lat[Math.min(lat.length - 1, Math.round(p / 100 * (lat.length - 1)))];
  console.log(JSON.stringify({
    requests: n, rate, seed, arrivals_s: reqs[n - 1].arrival, wall_s: wall, solves_per_s: n / wall,
    p50_s: pct(50), p95_s: pct(95), p99_s: pct(99), max_s: lat[lat.length - 1], converged, failed, iters_mean: iters / n,
    device_busy_s: s.gpuS, device_busy_frac: s.gpuS / wall, cell_updates_per_s: s.cellUpdates / wall,
    builds, criterion: `res/res0 <= ${resDrop} or ${maxIters} iterations, readback every 50 (the first takes the snapshot)`,
  }));
  srv.dispose();
  await pool.close();
}

if (require.main === module) main().catch(e => { console.error(e); process.exit(1); });
module.exports = { stream };
