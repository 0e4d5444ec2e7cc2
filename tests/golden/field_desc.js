'use strict';
/* Field model description written from a reference FieldModel object
 * (src/field.js:255-294): the constructor's products (stream knots after
 * assignMach, the processed exits, the inlet face velocity, the hot gas) and
 * the engine geometry FieldModel.sample reads (stream walls as their polyline
 * points, body polygons, combustor / mixer / afterburner, lip, outer line,
 * stations), as data. Used by make_golden.js and make_golden_axi.js. */
function fieldDesc(spec, m, cold) {
  const lp = L => L.pts.map(p => [p[0], p[1]]), c = m.c, fl = m.fl;
  const kn = {};
  for (const k of spec.order) kn[k] = m.kn[k].map(q => [q.x, q.Tt, q.Pt, q.M, q.sw, q.g, q.cp, q.R, q.mdot, q.noz]);
  const streams = {};
  for (const k of spec.order) { const st = spec.streams[k]; streams[k] = { x0: st.x0, x1: st.x1, lo: lp(st.lo), hi: lp(st.hi) }; }
  const cb = spec.comb, mx = spec.mixer, ab = spec.ab;
  const num = v => (typeof v === 'number' && isFinite(v) ? v : null);
  return {
    type: spec.type, s: spec.s, order: spec.order, streams, kn,
    bodies: m.bodies.map(b => ({ key: b.key, poly: b.poly.map(p => [p[0], p[1]]) })),
    c: { Tt3: num(c.Tt3), Tt4: num(c.Tt4), Tt5: num(c.Tt5), Tt6A: num(c.Tt6A), Tt7: num(c.Tt7), Tt13: num(c.Tt13), abOn: !!c.abOn },
    fl: { V0: fl.V0, T0: fl.T0, P0: fl.P0, Tt0: fl.Tt0, Pt0: fl.Pt0 },
    comb: { stream: cb.stream, x0: cb.x0, xd: cb.xd, x1: cb.x1, turbEnd: cb.turbEnd, n: cb.n, cannular: !!cb.cannular, linerIn: lp(cb.linerIn), linerOut: lp(cb.linerOut) },
    mixer: mx ? { x: mx.x, coreFrac: mx.coreFrac, lobes: mx.lobes } : null,
    ab: ab ? { stream: ab.stream, xFH: ab.xFH, x1: ab.x1, rings: ab.rings.slice() } : null,
    lip: { x: spec.lip.x, r: spec.lip.r, rOut: spec.lip.rOut }, outer: lp(spec.outer), outerEnd: spec.outerEnd,
    cold: { g: cold.g, cp: cold.cp },
    exits: m.exits.map(e => ({ key: e.key, x: e.x, R: e.R, Rin: e.Rin || 0, spread: e.spread, Qm: e.Qm, Qh: e.Qh, under: !!e.under, Ls: e.Ls, P: e.P, Pt: e.Pt, g: e.gas.g, cp: e.gas.cp, dTt: e.dTt })),
    Uface: m.Uface, maxdT: m.maxdT, hotGas: m.hotGas ? { g: m.hotGas.g, cp: m.hotGas.cp } : null,
    stations: spec.stations.map(s => ({ id: s.id, x: s.x })),
    view: { len: spec.view.len, Rmax: spec.view.Rmax }, domain: { xb: spec.domain.xb, rb: spec.domain.rb },
  };
}
module.exports = { fieldDesc };
