'use strict';
/* ============================================================================
 * cases.js — the workloads of this solver, built from their parameters.
 *
 *   sphereCase()      48 x 32 x 24 half domain, sphere of radius 4.8 cells on
 *                     the symmetry plane, M 0.4
 *   aircraftCase()    an aircraft of the stock family (airliner / interceptor /
 *                     fighter, or the 40 m transport) on a cfdGrid level
 *   seededCase()      one case of a seeded mix (batched and served forms)
 * ========================================================================== */
const path = require('path');
const fs = require('fs');
const B = require('./bodies.js');
const A = require('./air_cfd.js');

let SPECS = null;
/* Engine geometry records (nacelle meridian, lip, nozzle exits, casing radii)
 * of the three stock engines at their sea-level design point. */
function engineSpec(kind) {
  if (!SPECS) SPECS = JSON.parse(fs.readFileSync(path.join(__dirname, 'engine_specs.json'), 'utf8'));
  const s = SPECS[kind];
  if (!s) throw new Error(`no engine spec '${kind}'`);
  return JSON.parse(JSON.stringify(s));
}

/* Faces of a stretched axis with an exact cell count: a uniform core of
 * round((c1-c0)/h) cells and nlo / nhi cells growing geometrically by r. */
function stretchedFaces(c0, c1, h, nlo, nhi, r) {
  const n = Math.round((c1 - c0) / h), f = [];
  for (let i = 0; i <= n; i++) f.push(c0 + i * h);
  let d = h, x = c0;
  for (let i = 0; i < nlo; i++) { d *= r; x -= d; f.unshift(x); }
  d = h; x = c1;
  for (let i = 0; i < nhi; i++) { d *= r; x += d; f.push(x); }
  return f;
}

const SPHERE = { R: 4.8, M: 0.4, cells: [48, 32, 24] };
/* Sphere validation case: h = 1 core, radius 4.8 cells inside the core, x
 * stretched by 1.1, y by 1.06, z by 1.14; 48 x 32 x 24 cells. */
function sphereCase(R = SPHERE.R) {
  const fx = stretchedFaces(-6, 6, 1, 18, 18, 1.1);
  const fy = stretchedFaces(0, 6, 1, 0, 26, 1.06);
  const fz = stretchedFaces(-6, 6, 1, 6, 6, 1.14);
  const g = A.gridFromFaces(fx, fy, fz, 1, 'sphere');
  const body = B.sphereBody(R);
  const geo = A.cfdGeomBody(g, body);
  return { kind: 'sphere', g, geo, body, ref: { S: Math.PI * R * R, xcg: 0, zcg: 0, c: 1 } };
}

/* The sphere case refined or coarsened to about `cells` cells: every cell
 * count scales by f = cbrt(cells / 36864), the core spacing by 1/f and the
 * stretch ratios by their f-th root, so the physical box stays the same. */
function sphereCaseN(cells, R = SPHERE.R) {
  const f = Math.cbrt(cells / 36864), h = 1 / f;
  const n = (c) => Math.max(1, Math.round(c * f));
  const fx = stretchedFaces(-6, 6, h, n(18), n(18), Math.pow(1.1, 1 / f));
  const fy = stretchedFaces(0, 6, h, 0, n(26), Math.pow(1.06, 1 / f));
  const fz = stretchedFaces(-6, 6, h, n(6), n(6), Math.pow(1.14, 1 / f));
  const g = A.gridFromFaces(fx, fy, fz, h, 'sphere');
  const body = B.sphereBody(R);
  const geo = A.cfdGeomBody(g, body);
  return { kind: 'sphere', g, geo, body, ref: { S: Math.PI * R * R, xcg: 0, zcg: 0, c: 1 } };
}

/* Aircraft of the stock family on a grid of about `cells` cells, tail
 * deflection delta (deg). `over` overrides the parameter table. */
function aircraftCase(type, cells, deltaDeg = 0, over = null) {
  const T = B.AIRCRAFT_TYPES[type];
  const spec = engineSpec(T.engine);
  const ac = B.buildAircraft(type, over, spec);
  const G = new B.AirGeom(ac);
  const g = A.cfdGrid(ac, 'custom', cells);
  const geo = A.cfdGeom(ac, G, g, deltaDeg);
  return { kind: type, ac, G, g, geo, ref: ac.ref, xT: A.trefftzX(ac) };
}
/* The 40 m transport (airliner layout, AR 9.5, 25 deg sweep, taper 0.3, 12 %
 * sections, tail at 15 % of the wing area). */
function transportCase(cells, deltaDeg = 0) { return aircraftCase('airliner', cells, deltaDeg, B.TRANSPORT_OVERRIDES); }

/* Fixed states of the nacelle groups [inlet sink, bypass jet, core jet] from
 * (density, axial velocity, pressure) ratios to the freestream. */
function nacelleStates(sink, bypass, core) {
  const st = ([rr, uu, pr]) => { const pp = pr / A.CG; return [rr, rr * uu, 0, 0, pp / (A.CG - 1) + 0.5 * rr * uu * uu]; };
  return [st(sink), st(bypass), st(core)];
}
const TRANSPORT_FIX = nacelleStates([1.05, 0.45, 1.1], [0.95, 1.0, 1.3], [0.5, 1.6, 1.2]);

module.exports = { engineSpec, stretchedFaces, SPHERE, sphereCase, sphereCaseN, aircraftCase, transportCase, nacelleStates, TRANSPORT_FIX };
