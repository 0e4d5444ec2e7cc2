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
 *   const H = s.read();                    // {r, ru, rv, rw, E, dt}: Float32Array(N), reused buffers
 *   s.dispose();
 * Extras: setForceList(list), forces(alpha, M, S, xcg, zcg, cref), trefftz(xT, alpha, M, S),
 * residual(), autotune(iters), timeSteps(n).
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
  read() {
    const H = this.host;
    loadAddon().read(this.h, H.r, H.ru, H.rv, H.rw, H.E, H.dt);
    return H;
  }
  write(H) { loadAddon().write(this.h, H.r, H.ru, H.rv, H.rw, H.E); }
  sync() { loadAddon().sync(this.h); }
  setForceList(list) {
    const n = list.length, q = new Int32Array(n), cx = new Float64Array(n), cz = new Float64Array(n), x = new Float64Array(n), z = new Float64Array(n);
    for (let s = 0; s < n; s++) { const f = list[s]; q[s] = f.q; cx[s] = f.cx; cz[s] = f.cz; x[s] = f.x; z[s] = f.z; }
    loadAddon().setForceList(this.h, q, cx, cz, x, z);
  }
  forces(alphaDeg, M, S, xcg, zcg, cref) { return loadAddon().forces(this.h, alphaDeg, M, S, xcg, zcg, cref); }
  trefftz(xT, alphaDeg, M, S) { return loadAddon().trefftz(this.h, xT, alphaDeg, M, S); }
  residual() { return loadAddon().residual(this.h); }
  autotune(iters = 10) { this._syncOpt(); return loadAddon().autotune(this.h, iters); }
  timeSteps(n) { this._syncOpt(); const ms = loadAddon().timeSteps(this.h, n); this.it += n; return ms; }
  dispose() { if (this.h) { loadAddon().dispose(this.h); this.h = null; } }
}

module.exports = { CfdHIP, loadAddon };
