'use strict';
/* ============================================================================
 * geom_workers.js — the cut-cell build of an aircraft (cfdGeom,
 * src/air_cfd.js:50-141) on worker threads, with the masks of the one-thread
 * build byte for byte.
 *
 * The classification (corner flags, 27-point volume fractions, 9-point face
 * apertures) is most of the work and each z slab of it depends only on the
 * solid test, so the planes are cut into slabs that the workers and the
 * calling thread claim from a shared counter. Each slab computes its own
 * bounding corner planes, and every cell is written by exactly one slab, so the
 * values do not depend on the split. The baffles, fixed-state cells and face
 * cleanup follow on the calling thread (cfdGeom with the classification given).
 *
 *   const pool = new GeomPool(7);
 *   const C = await pool.aircraftCase('airliner', 1e6, 0, B.TRANSPORT_OVERRIDES);
 *   pool.close();
 * The result has the fields of cases.js aircraftCase. A worker keeps the last
 * aircraft it rebuilt, so a pool that serves several requests of one body
 * rebuilds it once. pool.axiSetup(spec, level) does the same for the
 * axisymmetric setup: buildZones' cut-cell pass (independent per grid
 * column) in column slabs, the remaining passes on the calling thread.
 * ========================================================================== */
const { Worker, isMainThread, parentPort } = require('worker_threads');
const A = require('./air_cfd.js');
const B = require('./bodies.js');
const K = require('./cases.js');
const Z = require('./cfd_grid.js');

// aircraft, solid test and grid of a recipe {type, cells, delta, over}
function rebuild(r) {
  const T = B.AIRCRAFT_TYPES[r.type];
  if (!T) throw new Error('geom_workers: unknown aircraft type ' + r.type);
  const ac = B.buildAircraft(r.type, r.over, K.engineSpec(T.engine));
  const G = new B.AirGeom(ac);
  const g = A.cfdGrid(ac, 'custom', r.cells);
  return { ac, G, g, solidAt: A.aircraftSolid(ac, G, g, r.delta).solidAt };
}

// claims slabs from the shared counter until none are left
function runSlabs(S, job) {
  const ctr = new Int32Array(job.ctr), n = job.nslab, NZ = S.g.NZ, len = job.slab;
  const part = { k0: 0, k1: 0, phi: new Float32Array(job.phi), ap: [0, 1, 2].map(d => new Float32Array(job.ap[d])) };
  let done = 0;
  for (;;) {
    const s = Atomics.add(ctr, 0, 1);
    if (s >= n) break;
    part.k0 = s * len; part.k1 = Math.min(NZ, part.k0 + len);
    A.classifyCells(S.g, S.solidAt, part);
    done++;
  }
  return done;
}

// engine description and grid of an axisymmetric setup {spec, level}
function rebuildAxi(r) {
  return { E: Z.engine(r.spec), grid: Z.buildCFDGrid(r.spec, r.level) };
}

// column slabs of buildZones' cut-cell pass, claimed from the shared counter
function runColumns(S, job) {
  const ctr = new Int32Array(job.ctr), W = S.grid.W, len = job.slab;
  const out = {
    phi: new Float32Array(job.bufs.phi), ae: new Float32Array(job.bufs.ae), an: new Float32Array(job.bufs.an),
    split: new Uint8Array(job.bufs.split), splitFr: new Float32Array(job.bufs.splitFr), stream: new Int8Array(job.bufs.stream),
    mat: new Uint8Array(job.bufs.mat),
  };
  let done = 0;
  for (;;) {
    const s = Atomics.add(ctr, 0, 1);
    if (s >= job.nslab) break;
    Z.zoneColumns(S.E, S.grid, out, s * len, Math.min(W, (s + 1) * len));
    done++;
  }
  return done;
}

if (!isMainThread) {
  let key = null, S = null;
  parentPort.on('message', job => {
    try {
      const k = JSON.stringify(job.recipe);
      if (k !== key) { S = job.axi ? rebuildAxi(job.recipe) : rebuild(job.recipe); key = k; }
      if ((job.axi ? S.grid.N : S.g.N) !== job.N) throw new Error('geom_workers: grid mismatch');
      parentPort.postMessage({ ok: true, slabs: job.axi ? runColumns(S, job) : runSlabs(S, job) });
    } catch (e) {
      parentPort.postMessage({ ok: false, error: String(e && e.stack || e) });
    }
  });
}

class GeomPool {
  constructor(n) {
    if (!(n >= 1)) throw new Error('GeomPool: at least one worker');
    this.workers = [];
    for (let w = 0; w < n; w++) this.workers.push(new Worker(__filename));
    this.busy = false;
  }

  /* cases.js aircraftCase(type, cells, deltaDeg, over) with the classification
   * spread over the pool; `slab` planes per claim (default: about four claims
   * per thread). */
  async aircraftCase(type, cells, deltaDeg = 0, over = null, slab = 0) {
    if (this.busy) throw new Error('GeomPool: one build at a time');
    if (!this.workers) throw new Error('GeomPool: closed');
    this.busy = true;
    try {
      const recipe = { type, cells, delta: deltaDeg || 0, over: over || null };
      const S = rebuild(recipe), g = S.g, N = g.N, NZ = g.NZ;
      const nt = this.workers.length + 1;
      const len = slab > 0 ? slab : Math.max(1, Math.ceil(NZ / (4 * nt)));
      const sab = () => new SharedArrayBuffer(4 * N);
      const job = { recipe, N, slab: len, nslab: Math.ceil(NZ / len), ctr: new SharedArrayBuffer(4), phi: sab(), ap: [sab(), sab(), sab()] };
      for (let d = 0; d < 3; d++) new Float32Array(job.ap[d]).fill(1);
      const replies = this.workers.map(w => new Promise(res => w.once('message', res)));
      for (const w of this.workers) w.postMessage(job);
      runSlabs(S, job);
      for (const r of await Promise.all(replies)) if (!r.ok) throw new Error(r.error);
      // plain (non-shared) arrays for the caller, like the one-thread build
      const pre = { phi: new Float32Array(job.phi).slice(), ap: [...job.ap.map(b => new Float32Array(b).slice()), new Float32Array(N)] };
      const geo = A.cfdGeom(S.ac, S.G, g, deltaDeg, pre);
      return { kind: type, ac: S.ac, G: S.G, g, geo, ref: S.ac.ref, xT: A.trefftzX(S.ac) };
    } finally {
      this.busy = false;
    }
  }

  transportCase(cells, deltaDeg = 0) { return this.aircraftCase('airliner', cells, deltaDeg, B.TRANSPORT_OVERRIDES); }

  /* cfd_grid.js axiSetup(spec, level) with buildZones' cut-cell pass (the
   * column pass, most of the work) spread over the pool: {grid, zones, comps},
   * the zone arrays byte-identical to the one-thread build. */
  async axiSetup(spec, level, slab = 0) {
    if (this.busy) throw new Error('GeomPool: one build at a time');
    if (!this.workers) throw new Error('GeomPool: closed');
    this.busy = true;
    try {
      const recipe = { spec, level: level || spec.level };
      const S = rebuildAxi(recipe), grid = S.grid, N = grid.N, W = grid.W;
      const nt = this.workers.length + 1;
      const len = slab > 0 ? slab : Math.max(1, Math.ceil(W / (4 * nt)));
      const bufs = {};
      for (const [k, b] of [['phi', 4], ['ae', 4], ['an', 4], ['split', 1], ['splitFr', 4], ['stream', 1], ['mat', 1]]) bufs[k] = new SharedArrayBuffer(b * N);
      const job = { axi: true, recipe, N, slab: len, nslab: Math.ceil(W / len), ctr: new SharedArrayBuffer(4), bufs };
      const replies = this.workers.map(w => new Promise(res => w.once('message', res)));
      for (const w of this.workers) w.postMessage(job);
      runColumns(S, job);
      for (const r of await Promise.all(replies)) if (!r.ok) throw new Error(r.error);
      // plain (non-shared) arrays for the caller, like the one-thread build
      const pre = {
        phi: new Float32Array(bufs.phi).slice(), ae: new Float32Array(bufs.ae).slice(), an: new Float32Array(bufs.an).slice(),
        split: new Uint8Array(bufs.split).slice(), splitFr: new Float32Array(bufs.splitFr).slice(), stream: new Int8Array(bufs.stream).slice(),
        mat: new Uint8Array(bufs.mat).slice(),
      };
      const zones = Z.buildZones(grid, spec, spec.op, pre);
      return { grid, zones, comps: Z.compParams(spec.type, spec.op, spec.eff) };
    } finally {
      this.busy = false;
    }
  }

  close() {
    if (!this.workers) return Promise.resolve();
    const w = this.workers; this.workers = null;
    return Promise.all(w.map(x => x.terminate()));
  }
}

module.exports = { GeomPool };
