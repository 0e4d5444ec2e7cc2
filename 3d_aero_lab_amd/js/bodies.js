'use strict';
/* ============================================================================
 * bodies.js — parametric bodies that the 3D cut-cell builder classifies.
 *
 * Three families:
 *   - the reference aircraft family (airliner / interceptor / fighter) with the
 *     same parameter table, fuselage, lifting-surface and inside predicates as
 *     the reference (src/aircraft.js:6-153, src/air_field.js:104-134), so that
 *     `cfdGeom` produces byte-identical cell flags, volume fractions and face
 *     apertures for the same parameters;
 *   - a sphere on the symmetry plane (the cut-cell validation case);
 *   - seeded variants of the aircraft family for batched and served solves.
 *
 * Bit-exact masks need V8 to evaluate the same floating-point operations in the
 * same order as the reference does, so every formula below keeps the operation
 * order of the reference expression it restates (cited per function). The
 * engine geometry (nacelle meridian, lip, nozzle exits) is data: the reference
 * derives it from its 1D cycle, which is outside this solver, and the values
 * for the three stock engines live in engine_specs.json.
 * ========================================================================== */

const DEG = Math.PI / 180;                       // src/field.js:7

/* Stock parameter table: src/aircraft.js:6-49 (geometry and the few mass/cg
 * fields the reference quantities use). */
const AIRCRAFT_TYPES = {
  airliner: {
    engine: 'hbtf', internal: false,
    p: {
      fusL: 37.6, fusD: 3.95, nose: 0.14, tail: 0.30, tailUp: 0.55, noseP: 0.5,
      wS: 122.6, wB: 35.8, wSweep: 27, wTaper: 0.24, wTcR: 0.15, wTcT: 0.11, wCamber: 0.02, wDih: 5, wInc: 2.0, wX: 0.34, wZ: -0.55,
      hS: 31, hB: 12.5, hSweep: 32, hTaper: 0.3, hTc: 0.10, hX: 0.80, hZ: 0.15,
      vS: 21.5, vH: 5.9, vSweep: 38, vTaper: 0.3, vTc: 0.10, vX: 0.78,
      engY: 0.34, cg: 0.25,
    },
  },
  interceptor: {
    engine: 'tj', internal: true,
    p: {
      fusL: 16.7, fusD: 1.55, nose: 0.28, tail: 0.14, tailUp: 0, noseP: 0.85,
      wS: 22, wB: 7.4, wSweep: 38, wTaper: 0.35, wTcR: 0.045, wTcT: 0.04, wCamber: 0, wDih: -4, wInc: 0, wX: 0.47, wZ: 0.0,
      hS: 5.5, hB: 3.9, hSweep: 38, hTaper: 0.35, hTc: 0.045, hX: 0.80, hZ: -0.45,
      vS: 4.6, vH: 2.7, vSweep: 42, vTaper: 0.35, vTc: 0.05, vX: 0.72,
      engY: 0, cg: 0.20,
    },
  },
  fighter: {
    engine: 'lbtf', internal: true,
    p: {
      fusL: 15.0, fusD: 1.45, nose: 0.26, tail: 0.13, tailUp: 0, noseP: 0.75,
      wS: 27.9, wB: 9.45, wSweep: 40, wTaper: 0.23, wTcR: 0.04, wTcT: 0.04, wCamber: 0, wDih: 0, wInc: 0, wX: 0.40, wZ: 0.0,
      hS: 5.9, hB: 5.6, hSweep: 40, hTaper: 0.3, hTc: 0.04, hX: 0.80, hZ: -0.4,
      vS: 5.1, vH: 2.6, vSweep: 47, vTaper: 0.4, vTc: 0.05, vX: 0.70,
      engY: 0, cg: 0.30,
    },
  },
};

/* NACA 4-digit half thickness with a closed trailing edge: src/aircraft.js:58. */
function nacaT(t, x) {
  const xx = Math.max(0, Math.min(1, x));
  return 5 * t * (0.2969 * Math.sqrt(xx) - 0.126 * xx - 0.3516 * xx * xx + 0.2843 * xx ** 3 - 0.1036 * xx ** 4);
}
/* NACA 4-digit mean camber line, maximum at chord fraction p: src/aircraft.js:59. */
function nacaC(m, x, p = 0.4) {
  if (!m) return 0;
  if (x < p) return m / (p * p) * (2 * p * x - x * x);
  return m / ((1 - p) ** 2) * ((1 - 2 * p) + 2 * p * x - x * x);
}

/* Body of revolution: power-law nose, cylinder, cosine boat-tail with an
 * upswept centreline (src/aircraft.js:63-75). */
function makeFuselage(p, endR) {
  const L = p.fusL, R0 = p.fusD / 2, Ln = p.nose * L, Lt = p.tail * L, xt0 = L - Lt;
  const Re = Math.min(endR, R0 * 0.95);
  function R(x) {
    if (x <= 0) return 0;
    if (x < Ln) { const s = 1 - x / Ln; return R0 * Math.pow(Math.max(0, 1 - s * s), p.noseP); }
    if (x <= xt0) return R0;
    if (x >= L) return Re;
    const t = (x - xt0) / Lt;
    return Re + (R0 - Re) * 0.5 * (1 + Math.cos(Math.PI * t));
  }
  function zc(x) { return x <= xt0 ? 0 : p.tailUp * R0 * ((x - xt0) / Lt) ** 2; }
  return { L, R0, Ln, Lt, xt0, Re, R, zc };
}

/* Trapezoidal lifting surface (wing, horizontal tail, or a fin on the
 * symmetry plane when o.vertical): src/aircraft.js:78-102. Only the members the
 * cut-cell builder, the grid box and the reference quantities read are built. */
function makeSurface(o) {
  const vertical = !!o.vertical;
  const semi = vertical ? o.b : o.b / 2;
  const cr = 2 * o.S / (o.b * (1 + o.taper));
  const ct = cr * o.taper;
  const tanL = Math.tan(o.sweep * Math.PI / 180);
  const tanD = Math.tan((o.dih || 0) * Math.PI / 180);
  const tanI = Math.tan((o.inc || 0) * Math.PI / 180);
  const chord = s => cr + (ct - cr) * s / semi;
  const xLE = s => o.x0 + s * tanL;
  const tc = s => o.tcR + (o.tcT - o.tcR) * s / semi;
  const mac = (2 / 3) * cr * (1 + o.taper + o.taper ** 2) / (1 + o.taper);
  const sMac = semi / 3 * (1 + 2 * o.taper) / (1 + o.taper);
  /* surface point at span station s, chord fraction xc, upper (+1) / lower (-1) side */
  const pt = (s, xc, side, sAtt, tipZero) => {
    const c = chord(s), x = xLE(s) + xc * c;
    let th = nacaT(tc(s), xc) * c;
    if (tipZero) th = 0;
    const cam = nacaC(o.camber || 0, xc) * c;
    const shear = -(xc - 0.25) * c * tanI;
    if (vertical) return [x, side * th, o.z0 + s + 0 * cam];
    return [x, s, o.z0 + (s - (sAtt || 0)) * tanD + cam + shear + side * th];
  };
  return { o, vertical, semi, cr, ct, tanL, chord, xLE, tc, mac, sMac, xMac: xLE(sMac), pt, S: o.S };
}

/* Aircraft assembly: src/aircraft.js:108-153. `spec` is an engine record from
 * engine_specs.json (or a synthetic one with the same fields). */
function buildAircraft(type, over, spec) {
  const T = AIRCRAFT_TYPES[type];
  if (!T) throw new Error(`unknown aircraft type '${type}'`);
  const p = Object.assign({}, T.p, over || {});
  const internal = over && over.internal !== undefined ? !!over.internal : T.internal;
  let endR = 0.06 * p.fusD / 2;
  if (internal) {
    p.fusD = Math.max(p.fusD, 2 * (spec.outerMaxR + 0.06));
    endR = spec.outerAtEnd + 0.01;
  }
  const fus = makeFuselage(p, endR);
  const R0 = fus.R0;
  const wing = makeSurface({ S: p.wS, b: p.wB, sweep: p.wSweep, taper: p.wTaper, tcR: p.wTcR, tcT: p.wTcT, camber: p.wCamber, dih: p.wDih, inc: p.wInc, x0: p.wX * fus.L, z0: p.wZ * R0 });
  const hx0 = p.hX * fus.L, hz0 = fus.zc(hx0 + 1) + p.hZ * fus.R(hx0 + 1);
  const htail = makeSurface({ S: p.hS, b: p.hB, sweep: p.hSweep, taper: p.hTaper, tcR: p.hTc, tcT: p.hTc, camber: 0, dih: 0, inc: 0, x0: hx0, z0: hz0 });
  const vx0 = p.vX * fus.L;
  const vtail = makeSurface({ S: p.vS, b: p.vH, sweep: p.vSweep, taper: p.vTaper, tcR: p.vTc, tcT: p.vTc, vertical: true, x0: vx0, z0: 0 });
  const vxm = vx0 + 0.5 * vtail.cr;
  vtail.o.z0 = fus.zc(vxm) + fus.R(vxm) * 0.85;           // fin root on the fuselage crown
  const engines = [];
  if (internal) {
    engines.push({ x0: fus.L - spec.outerEnd, y: 0, z: 0, s: 1, internal: true });
  } else {
    const ye = p.engY * wing.semi, Lnac = spec.outerEnd, Rn = spec.view.Rmax;
    const xLip = wing.xLE(ye) - Lnac + 0.35 * wing.chord(ye) * 0.25;
    const zw = wing.pt(ye, 0.02, -1, R0)[2];
    engines.push({ x0: xLip, y: ye, z: zw - Rn - 0.28, s: 1, internal: false, Lnac, Rn,
      meridian: spec.meridian ? spec.meridian.map(q => q.slice()) : null });
  }
  const xcg = wing.xMac + p.cg * wing.mac;
  return { type, p, fus, wing, htail, vtail, engines, spec,
    ref: { S: p.wS, b: p.wB, c: wing.mac, A: p.wB * p.wB / p.wS, xcg, zcg: 0, xMac: wing.xMac, yMac: wing.sMac } };
}

/* Inside tests for the aircraft: src/air_field.js:104-134. */
class AirGeom {
  constructor(ac) {
    this.ac = ac;
    const fus = ac.fus;
    const rootSpan = S => Math.min(fus.R((S.xLE(fus.R0) * 2 + S.chord(fus.R0)) / 2) * 0.98, S.semi * 0.5);
    const box = S => {
      let x0 = Infinity, x1 = -Infinity, z0 = Infinity, z1 = -Infinity;
      for (let k = 0; k <= 10; k++) {
        const s = S.semi * k / 10;
        for (const xc of [0, 1]) for (const sd of [-1, 1]) {
          const q = S.pt(s, xc, sd, 0);
          x0 = Math.min(x0, q[0]); x1 = Math.max(x1, q[0]); z0 = Math.min(z0, q[2]); z1 = Math.max(z1, q[2]);
        }
      }
      return [x0, x1, z0 - 0.2 * S.cr, z1 + 0.2 * S.cr];
    };
    this.sW = rootSpan(ac.wing); this.sH = rootSpan(ac.htail);
    this.bW = box(ac.wing); this.bH = box(ac.htail);
  }
  inLift(S, sA, bx, x, ay, z) {
    if (x < bx[0] || x > bx[1] || z < bx[2] || z > bx[3] || ay > S.semi) return false;
    const c = S.chord(ay), xc = (x - S.xLE(ay)) / c;
    if (xc < 0 || xc > 1) return false;
    const th = nacaT(S.tc(ay), xc) * c * (ay >= S.semi - 1e-6 ? 0 : 1);
    const cam = nacaC(S.o.camber || 0, xc) * c, shear = -(xc - 0.25) * c * Math.tan((S.o.inc || 0) * DEG);
    const zm = S.o.z0 + (ay - sA) * Math.tan((S.o.dih || 0) * DEG) + cam + shear;
    return z >= zm - th && z <= zm + th;
  }
  inVt(x, ay, z) {
    const V = this.ac.vtail, s = z - V.o.z0;
    if (s < -0.5 || s > V.semi) return false;
    const ss = Math.max(0, s), c = V.chord(ss), xc = (x - V.xLE(ss)) / c;
    if (xc < 0 || xc > 1) return false;
    return ay <= nacaT(V.tc(ss), xc) * c;
  }
  inFus(x, ay, z) {
    const f = this.ac.fus;
    if (x < 0 || x > f.L) return false;
    const R = f.R(x), dz = z - f.zc(x);
    return ay * ay + dz * dz <= R * R;
  }
}

/* ---------------------------------------------------------------------------
 * Sphere on the symmetry plane. The cut-cell builder only needs a predicate
 * (strict inside, |y| mirrored), so the sphere is a body with no lifting
 * surfaces, baffles or engines.
 * ------------------------------------------------------------------------- */
function sphereBody(R, cx = 0, cz = 0) {
  const R2 = R * R;
  return {
    kind: 'sphere', R, cx, cz,
    solidAt(x, y, z) { const dx = x - cx, dz = z - cz; return dx * dx + y * y + dz * dz < R2; },
  };
}

/* mulberry32: 32-bit seeded generator used for every seeded case so that the
 * JS host and the C++ drivers draw the same sequence. */
function mulberry32(seed) {
  let a = seed >>> 0;
  return function () {
    a = (a + 0x6D2B79F5) >>> 0;
    let t = a;
    t = Math.imul(t ^ (t >>> 15), t | 1);
    t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}

/* Transport of the plain 1.0M-cell workload: the airliner layout with a 40 m x
 * 4 m fuselage, aspect-ratio 9.5 wing swept 25 deg with taper 0.3 and 12 %
 * sections, horizontal tail at 15 % of the wing area. */
const TRANSPORT_OVERRIDES = (() => {
  const wS = 122.6;
  return { fusL: 40, fusD: 4, wS, wB: Math.sqrt(9.5 * wS), wSweep: 25, wTaper: 0.3, wTcR: 0.12, wTcT: 0.12, hS: 0.15 * wS };
})();

if (typeof module !== 'undefined') module.exports = {
  DEG, AIRCRAFT_TYPES, nacaT, nacaC, makeFuselage, makeSurface, buildAircraft, AirGeom, sphereBody, mulberry32,
  TRANSPORT_OVERRIDES,
};
