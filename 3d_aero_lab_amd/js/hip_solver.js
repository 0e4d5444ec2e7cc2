'use strict';
/* ============================================================================
 * hip_solver.js — CfdHIP: the solver object of the reference's protocol
 * (CfdCPU / CfdGPU, src/air_cfd.js:157-259, 367-438) on one MI355X through
 * the Node-API addon (lib/aero3d.node) over libaero3d.so.
 *
 *   const s = new CfdHIP(g, geo);          // throws Error when no device / bad grid
 *   s.opt.cfl = 0.5; s.opt.order = 2;      // read at every step
 *   s.setFree(F); s.setFix([f3, f4, f5]); s.reset();
 *   s.step(n);                             // enqueued on the solver's stream
 *   await s.stepAsync(n);                  // the same, settled once the iterations ran
 *   const H = s.read();                    // {r, ru, rv, rw, E, dt}: Float32Array(N), reused buffers
 *   s.dispose();
 * Extras: setForceList(list), forces(alpha, M, S, xcg, zcg, cref), trefftz(xT, alpha, M, S),
 * residual(), readback(ref, xT), probe(pts, fs, M), autotune(iters), timeSteps(n).
 *
 * CfdMonitor: the readback controller state of the reference (CFD3.readback,
 * src/air_cfd.js:550-578) over a solver's device readback: residual, force
 * coefficients, Trefftz CDi, Mach maximum, history and the convergence flag
 * (ConvRule). Batch / Server: the batched and served forms (many solves in
 * one device arena; continuous batching with on-device convergence).
 * ========================================================================== */
const path = require('path');

let addon = null;
function loadAddon() {
  if (!addon) addon = require(path.join(__dirname, '..', 'lib', 'aero3d.node'));
  return addon;
}

class CfdHIP {
  constructor(g, geo) {
    const A = loadAddon();
    this.kind = 'hip'; this.g = g; this.geo = geo; this.t = geo.t; this.it = 0;
    this.opt = { cfl: 0.7, order: 2 };                     // src/air_cfd.js:162
    this.h = A.create(Float64Array.from(g.X.f), Float64Array.from(g.Y.f), Float64Array.from(g.Z.f), geo.t, geo.phi,
      geo.ap[0], geo.ap[1], geo.ap[2], geo.ap[3]);
    const N = g.N;
    this.host = { r: new Float32Array(N), ru: new Float32Array(N), rv: new Float32Array(N), rw: new Float32Array(N),
      E: new Float32Array(N), dt: new Float32Array(N) };
    this._opt = null;
  }
  _syncOpt() {
    const o = this.opt.cfl + ',' + this.opt.order;
    if (o !== this._opt) { loadAddon().setOpt(this.h, this.opt.cfl, this.opt.order); this._opt = o; }
  }
  setFree(F) { this.F = F.slice(); loadAddon().setFree(this.h, Float64Array.from(F)); }
  setFix(fx) { this.fix = fx.map(a => a.slice()); loadAddon().setFix(this.h, Float64Array.from([].concat(...fx))); }
  reset() { this._syncOpt(); loadAddon().reset(this.h); this.it = 0; }
  step(n) { this._syncOpt(); loadAddon().step(this.h, n); this.it += n; }
  /* n iterations on a worker thread; the promise settles once they have run. The
   * handle is busy meanwhile (any other call on it throws). */
  stepAsync(n) { this._syncOpt(); const p = loadAddon().stepAsync(this.h, n); this.it += n; return p; }
  read() {
    const H = this.host;
    loadAddon().read(this.h, H.r, H.ru, H.rv, H.rw, H.E, H.dt);
    return H;
  }
  write(H) { loadAddon().write(this.h, H.r, H.ru, H.rv, H.rw, H.E); }
  sync() { loadAddon().sync(this.h); }
  setForceList(list) { const f = forceArrays(list); loadAddon().setForceList(this.h, f.q, f.cx, f.cz, f.x, f.z); }
  forces(alphaDeg, M, S, xcg, zcg, cref) { return loadAddon().forces(this.h, alphaDeg, M, S, xcg, zcg, cref); }
  trefftz(xT, alphaDeg, M, S) { return loadAddon().trefftz(this.h, xT, alphaDeg, M, S); }
  residual() { return loadAddon().residual(this.h); }
  /* CFD3.readback's numbers on the device in one call: ref = {alpha, M, S, xcg, zcg, cref};
   * xT: Trefftz plane x (NaN or undefined: none). M is floored at 0.05 as the reference does. */
  readback(ref, xT) {
    return loadAddon().readback(this.h, ref.alpha || 0, ref.M, ref.S, ref.xcg || 0, ref.zcg || 0, ref.cref || 1,
      xT === undefined ? NaN : xT);
  }
  /* Display sampling of the current state on the device (the reference's
   * cfdStencil / cfdSample / cfdState / cfdVelAt, src/air_cfd.js:453-490, 637-646)
   * at the points pts = [x0, y0, z0, x1, ...] of the full model. fs: the
   * freestream record {T, a, P, rho, q} of the dimensional state (null: none);
   * M: the freestream Mach number of the velocities. Returns {n, idx (8n), w (8n),
   * ok (n), W (5n: rho, u, v, w, p; NaN where !ok), state (12n, PROBE_KEYS order,
   * or null), vel (3n, in units of the freestream speed; NaN where !vok), vok (n)}. */
  probe(pts, fs, M) {
    const P = pts instanceof Float64Array ? pts : Float64Array.from(pts), n = P.length / 3;
    const o = { n, idx: new Int32Array(8 * n), w: new Float32Array(8 * n), ok: new Uint8Array(n), W: new Float64Array(5 * n),
      state: fs ? new Float64Array(12 * n) : null, vel: new Float64Array(3 * n), vok: new Uint8Array(n) };
    loadAddon().probe(this.h, P, fs ? Float64Array.of(fs.T, fs.a, fs.P, fs.rho, fs.q) : null, M || 0, o.idx, o.w, o.ok, o.W,
      o.state, o.vel, o.vok);
    return o;
  }
  autotune(iters = 10) { this._syncOpt(); return loadAddon().autotune(this.h, iters); }
  timeSteps(n) { this._syncOpt(); const ms = loadAddon().timeSteps(this.h, n); this.it += n; return ms; }
  dispose() { if (this.h) { loadAddon().dispose(this.h); this.h = null; } }
}

/* columns of CfdHIP.probe's state, the fields of the reference's cfdState */
const PROBE_KEYS = ['P', 'T', 'rho', 'V', 'M', 'Tt', 'Pt', 'Cp', 'Taw', 'Vx', 'Vy', 'Vz'];

function forceArrays(list) {
  const n = list.length, q = new Int32Array(n), cx = new Float64Array(n), cz = new Float64Array(n), x = new Float64Array(n), z = new Float64Array(n);
  for (let s = 0; s < n; s++) { const f = list[s]; q[s] = f.q; cx[s] = f.cx; cz[s] = f.cz; x[s] = f.x; z[s] = f.z; }
  return { q, cx, cz, x, z };
}

/* Convergence rules, one update per readback; the same rules the served form
 * decides on the device (csrc/converge.hpp).
 *   RULE_REFERENCE  the reference controller (src/air_cfd.js:573-576): more than six
 *                   readbacks, res / res0 < 2e-3 and the lift coefficients of the last
 *                   four readbacks within 2e-3; done then or past `cap` iterations
 *   RULE_RES_DROP   res / res0 <= resDrop; done then or at `cap` iterations
 * res0 is the first positive residual. Both end a solve whose residual is not
 * finite after its first readback (diverged). */
const RULE_RES_DROP = 0, RULE_REFERENCE = 1;
class ConvRule {
  constructor(rule = RULE_REFERENCE, resDrop = 1e-4, cap = 30000) {
    this.rule = rule; this.resDrop = resDrop; this.cap = cap;
    this.res0 = NaN; this.cl = []; this.nreads = 0; this.conv = false; this.diverged = false; this.done = false;
  }
  update(res, CL, it) {
    this.nreads++;
    if (!(this.res0 > 0) && res > 0) this.res0 = res;
    this.cl.push(CL); if (this.cl.length > 4) this.cl.shift();
    this.diverged = this.nreads > 1 && !Number.isFinite(res);
    if (this.rule === RULE_REFERENCE) {
      let ok = this.nreads > 6 && this.res0 > 0 && res / this.res0 < 2e-3;
      if (ok) {
        let lo = Infinity, hi = -Infinity, nan = false;
        for (const c of this.cl) { if (c !== c) nan = true; if (c < lo) lo = c; if (c > hi) hi = c; }
        ok = !nan && hi - lo < 2e-3;
      }
      this.conv = ok;
      this.done = this.conv || this.diverged || it > this.cap;
    } else {
      this.conv = this.res0 > 0 && res / this.res0 <= this.resDrop;
      this.done = this.conv || this.diverged || it >= this.cap;
    }
    return this.done;
  }
}

/* The readback controller state of the reference (CFD3.readback,
 * src/air_cfd.js:550-578) over a solver's device readback: the first read
 * takes the density snapshot, every later one the residual since the previous
 * read. ref = {alpha, M, S, xcg, zcg, cref}; xT: Trefftz plane (optional). */
class CfdMonitor {
  constructor(solver, ref, xT, rule = new ConvRule()) {
    this.s = solver; this.ref = ref; this.xT = xT; this.rule = rule;
    this.hist = []; this.res = NaN; this.res0 = NaN; this.conv = false; this.F = null; this.CDi = NaN; this.Mmax = NaN;
  }
  read() {
    const o = this.s.readback(this.ref, this.xT);
    this.res = o.res; this.F = { CL: o.CL, CD: o.CD, Cm: o.Cm }; this.CDi = o.CDi; this.Mmax = o.Mmax;
    this.rule.update(o.res, o.CL, o.it);
    this.res0 = this.rule.res0; this.conv = this.rule.conv;
    this.hist.push({ it: o.it, res: o.res, CL: o.CL, CD: o.CD, Cm: o.Cm });
    if (this.hist.length > 400) this.hist.splice(0, this.hist.length - 400);
    return o;
  }
  /* Step to the rule's end: a read every `every` iterations (and at the cap of
   * the residual-drop rule), the first taking the density snapshot only, as
   * the reference controller's first readback does; returns the last readback. */
  solve(every = 50) {
    let o = null;
    while (!this.rule.done) {
      let nxt = (Math.floor(this.s.it / every) + 1) * every;
      if (this.rule.rule === RULE_RES_DROP) nxt = Math.min(nxt, this.rule.cap);
      this.s.step(nxt - this.s.it);
      o = this.read();
    }
    return o;
  }
}

function geometryArgs(g, geo, list) {
  const f = forceArrays(list || []);
  return [Float64Array.from(g.X.f), Float64Array.from(g.Y.f), Float64Array.from(g.Z.f), geo.t, geo.phi, geo.ap[0], geo.ap[1], geo.ap[2],
    geo.ap[3], f.q, f.cx, f.cz, f.x, f.z];
}

/* Batched form: many solves of uneven grid sizes in one device arena, stepped
 * together (one launch per RK stage over all cases, graph-replayed). */
class Batch {
  constructor(arenaBytes = 8 * 2 ** 30, maxSlots = 256) { this.h = loadAddon().batchCreate(arenaBytes, maxSlots); this.N = new Map(); }
  addGeometry(g, geo, list) { const gid = loadAddon().batchAddGeometry(this.h, ...geometryArgs(g, geo, list)); this.N.set(gid, g.N); return gid; }
  removeGeometry(gid) { loadAddon().batchRemoveGeometry(this.h, gid); }
  add(gid, F, fix, cfl = 0.5, order = 2) {
    const slot = loadAddon().batchAdd(this.h, gid, Float64Array.from(F), Float64Array.from([].concat(...fix)), cfl, order);
    this.N.set(-1 - slot, this.N.get(gid));
    return slot;
  }
  remove(slot) { loadAddon().batchRemove(this.h, slot); }
  step(n) { loadAddon().batchStep(this.h, n); }
  stepAsync(n) { return loadAddon().batchStepAsync(this.h, n); }
  sync() { loadAddon().batchSync(this.h); }
  iter(slot) { return loadAddon().batchIter(this.h, slot); }
  read(slot) {
    const N = this.N.get(-1 - slot), H = { r: new Float32Array(N), ru: new Float32Array(N), rv: new Float32Array(N), rw: new Float32Array(N), E: new Float32Array(N) };
    loadAddon().batchRead(this.h, slot, H.r, H.ru, H.rv, H.rw, H.E);
    return H;
  }
  residuals(slots) { const out = new Float64Array(slots.length); loadAddon().batchResiduals(this.h, Int32Array.from(slots), out); return out; }
  forces(slot, ref) { return loadAddon().batchForces(this.h, slot, ref.alpha || 0, ref.M, ref.S, ref.xcg || 0, ref.zcg || 0, ref.cref || 1); }
  setOption(name, value) { loadAddon().batchSetOption(this.h, name, value); }
  footprint() { return loadAddon().batchFootprint(this.h); }
  dispose() { if (this.h) { loadAddon().batchDispose(this.h); this.h = null; } }
}

/* Served form: requests queue FIFO and are admitted by device-memory
 * footprint (continuous batching); every tick steps all resident solves to the
 * next readback of any of them, and each solve's rule is decided on the
 * device. request = {geometry, F, fix, cfl, order, alpha, M, S, xcg, zcg, cref,
 * rule, resDrop, maxIters, arrival, tag}. */
class Server {
  constructor(arenaBytes = 8 * 2 ** 30, maxSlots = 256, checkEvery = 50) { this.h = loadAddon().serverCreate(arenaBytes, maxSlots, checkEvery); }
  addGeometry(g, geo, list) { return loadAddon().serverAddGeometry(this.h, ...geometryArgs(g, geo, list)); }
  removeGeometry(gid) { loadAddon().serverRemoveGeometry(this.h, gid); }
  submit(r) {
    loadAddon().serverSubmit(this.h, Object.assign({}, r, { F: Float64Array.from(r.F), fix: Float64Array.from([].concat(...r.fix)) }));
  }
  tick() { return loadAddon().serverTick(this.h); }
  tickAsync() { return loadAddon().serverTickAsync(this.h); }
  poll() { return loadAddon().serverPoll(this.h); }
  clock() { return loadAddon().serverClock(this.h); }
  stats() { return loadAddon().serverStats(this.h); }
  dispose() { if (this.h) { loadAddon().serverDispose(this.h); this.h = null; } }
}

module.exports = { CfdHIP, CfdMonitor, ConvRule, RULE_RES_DROP, RULE_REFERENCE, Batch, Server, forceArrays, loadAddon, PROBE_KEYS };
