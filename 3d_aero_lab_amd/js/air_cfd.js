'use strict';
/* ============================================================================
 * air_cfd.js — JS host layer of the MI355X 3D Euler cut-cell solver.
 *
 * Drop-in for the reference entry points of src/air_cfd.js:
 *   cfdAxis / cfdGrid        stretched Cartesian grid        (src/air_cfd.js:14-44)
 *   cfdGeom / cfdForceList   cut-cell masks, force stencil   (src/air_cfd.js:50-152)
 *   CfdHIP                   solver object with the CfdCPU / CfdGPU protocol
 *                            (setFree, setFix, reset, step, read, dispose; opt)
 *                            backed by the HIP library      (src/air_cfd.js:157-259, 367-438)
 * The post-processing the reference runs on read() output (cfdForces, the
 * readback's residual / Trefftz sums, display sampling) runs on the device:
 * CfdHIP.forces / trefftz / residual / readback / probe of hip_solver.js.
 *
 * Grid and mask construction stays in JS doubles: the masks are compared
 * byte-for-byte with the reference, so every expression that feeds a
 * classification keeps the operation order of the reference expression.
 * ========================================================================== */

const { DEG, nacaC } = require('./bodies.js');

const CG = 1.4;                                                   // gamma
const CFD_GRIDS = { coarse: { cells: 150e3 }, medium: { cells: 380e3 }, fine: { cells: 820e3 } };  // src/air_cfd.js:10
const SUB = [1 / 6, 0.5, 5 / 6];                                  // 3-point sub-sampling per axis

/* One grid axis: a uniform core of round((c1-c0)/h) cells over [c0,c1] and a
 * geometric stretch by ratio r out to lo / hi; the last stretched face is
 * snapped to the boundary and a sliver cell merged into its neighbour
 * (src/air_cfd.js:14-20). Returns face coordinates. */
function cfdAxis(c0, c1, h, lo, hi, r) {
  const n = Math.max(1, Math.round((c1 - c0) / h)), hh = (c1 - c0) / n;
  const f = new Array(n + 1);
  for (let i = 0; i <= n; i++) f[i] = c0 + i * hh;
  if (lo < c0 - 1e-9) {
    const L = [];
    for (let d = hh, x = c0; x > lo + 1e-9;) { d *= r; x -= d; L.push(x); }
    const m = L.length;
    L[m - 1] = lo;
    if (m > 1 && (L[m - 2] - lo) < 0.5 * (m > 2 ? L[m - 3] - L[m - 2] : hh)) L.splice(m - 2, 1);
    L.reverse();
    f.unshift(...L);
  }
  if (hi > c1 + 1e-9) {
    const R = [];
    for (let d = hh, x = c1; x < hi - 1e-9;) { d *= r; x += d; R.push(x); }
    const m = R.length;
    R[m - 1] = hi;
    if (m > 1 && (hi - R[m - 2]) < 0.5 * (m > 2 ? R[m - 2] - R[m - 3] : hh)) R.splice(m - 2, 1);
    f.push(...R);
  }
  return f;
}

/* Axis record from faces: Float64 faces, Float32 centres and widths
 * (src/air_cfd.js:36). */
function axisFromFaces(f) {
  const n = f.length - 1, c = new Float32Array(n), d = new Float32Array(n);
  for (let i = 0; i < n; i++) { c[i] = 0.5 * (f[i] + f[i + 1]); d[i] = f[i + 1] - f[i]; }
  return { f: Float64Array.from(f), c, d, n };
}
function gridFromFaces(fx, fy, fz, h, level) {
  const X = axisFromFaces(fx), Y = axisFromFaces(fy), Z = axisFromFaces(fz);
  return { X, Y, Z, NX: X.n, NY: Y.n, NZ: Z.n, N: X.n * Y.n * Z.n, h, level: level || 'custom' };
}

/* Half-domain grid around an aircraft: body box with margins, growth ratio
 * 1.14, core spacing h iterated until the count lands within 4 % of the
 * level's target (src/air_cfd.js:21-44). `cells` overrides the level target. */
function cfdGrid(ac, level, cells) {
  const L = ac.fus.L, w = ac.wing, ht = ac.htail, vt = ac.vtail, sp = ac.spec, Lc = Math.max(L, ac.ref.b);
  let x0 = -0.02 * L;
  let x1 = Math.max(L, ht.xLE(ht.semi) + ht.ct, vt.xLE(vt.semi) + vt.ct) + 0.03 * L;
  let y1 = Math.max(w.semi, ht.semi) + 0.03 * Lc;
  let z0 = -ac.fus.R0 - 0.03 * L;
  const z1 = vt.o.z0 + vt.semi + 0.02 * L;
  const xLast = sp.exitsGeom[sp.exitsGeom.length - 1].x;
  for (const e of ac.engines) {
    x0 = Math.min(x0, e.x0 - 0.02 * L);
    x1 = Math.max(x1, e.x0 + xLast + 0.05 * L);
    if (!e.internal) { y1 = Math.max(y1, e.y + sp.view.Rmax + 0.03 * Lc); z0 = Math.min(z0, e.z - sp.view.Rmax - 0.03 * L); }
  }
  z0 = Math.min(z0, w.pt(w.semi, 0.5, 0, 0)[2] - 0.05 * L, ht.o.z0 - 0.05 * L);
  const lim = { x0: x0 - 0.55 * Lc, x1: x1 + 0.9 * Lc, y1: y1 + 0.6 * Lc, z0: z0 - 0.6 * Lc, z1: z1 + 0.6 * Lc };
  const target = cells || CFD_GRIDS[level].cells, r = 1.14;
  const axes = hh => [cfdAxis(x0, x1, hh, lim.x0, lim.x1, r), cfdAxis(0, y1, hh, 0, lim.y1, r), cfdAxis(z0, z1, hh, lim.z0, lim.z1, r)];
  let h = Math.cbrt((x1 - x0) * y1 * (z1 - z0) / (target * 0.6));
  for (let it = 0; it < 6; it++) {
    const ax = axes(h);
    const n = (ax[0].length - 1) * (ax[1].length - 1) * (ax[2].length - 1);
    h *= Math.cbrt(n / target);
    if (Math.abs(n / target - 1) < 0.04) break;
  }
  const ax = axes(h);
  const g = gridFromFaces(ax[0], ax[1], ax[2], h, level);
  g.core = { x0, x1, y1, z0, z1 };
  return g;
}
const cidx = (g, i, j, k) => (k * g.NY + j) * g.NX + i;              // src/air_cfd.js:45

/* ---------------------------------------------------------------------------
 * Cut-cell classification shared by every body.
 *   corner flags -> volume fraction phi (27 sub-points in mixed cells)
 *   -> + face apertures ap[0..2] (9 sub-points on mixed faces)
 * (src/air_cfd.js:68-90). phi and ap are exact multiples of 1/27 and 1/9.
 * ------------------------------------------------------------------------- */
function classifyCells(g, solidAt, part) {
  const NX = g.NX, NY = g.NY, NZ = g.NZ, N = g.N, X = g.X, Y = g.Y, Z = g.Z;
  const sx = NX + 1, sxy = sx * (NY + 1);
  // planes k0..k1-1 (all by default) into part.phi / part.ap when given; the
  // corner planes k0..k1 are computed locally, so slabs are independent
  const k0 = part ? part.k0 : 0, k1 = part ? part.k1 : NZ;
  const corner = new Uint8Array(sxy * (k1 - k0 + 1));
  // one plane per call: the engine compiles the per-plane loop after the first
  // few planes instead of interpreting one long loop
  const cornerPlane = k => {
    const o = (k - k0) * sxy;
    for (let j = 0; j <= NY; j++) for (let i = 0; i <= NX; i++) corner[o + j * sx + i] = solidAt(X.f[i], Y.f[j], Z.f[k]) ? 1 : 0;
  };
  for (let k = k0; k <= k1; k++) cornerPlane(k);
  const phi = part ? part.phi : new Float32Array(N);
  const ap = part ? part.ap : [new Float32Array(N).fill(1), new Float32Array(N).fill(1), new Float32Array(N).fill(1), new Float32Array(N)];
  const cellPlane = k => { for (let j = 0; j < NY; j++) for (let i = 0; i < NX; i++) cell(i, j, k); };
  const cell = (i, j, k) => {
    const q = (k * NY + j) * NX + i;
    // corners (i + a, j + b, k + c) at b0 + a + b sx + c sxy
    const b0 = (k - k0) * sxy + j * sx + i, b1 = b0 + sxy;
    const c0 = corner[b0], c1 = corner[b0 + 1], c2 = corner[b0 + sx], c3 = corner[b0 + sx + 1];
    const c4 = corner[b1], c5 = corner[b1 + 1], c6 = corner[b1 + sx], c7 = corner[b1 + sx + 1];
    const s = c0 + c1 + c2 + c3 + c4 + c5 + c6 + c7;
    if (s === 0) phi[q] = 1;
    else if (s === 8) phi[q] = 0;
    else {
      let n = 0;
      for (const a of SUB) for (const b of SUB) for (const c of SUB)
        if (!solidAt(X.f[i] + a * X.d[i], Y.f[j] + b * Y.d[j], Z.f[k] + c * Z.d[k])) n++;
      phi[q] = n / 27;
    }
    for (let d = 0; d < 3; d++) {
      // the four corners of the + face normal to axis d
      let m;
      if (d === 0) m = c1 + c3 + c5 + c7;
      else if (d === 1) m = c2 + c3 + c6 + c7;
      else m = c4 + c5 + c6 + c7;
      if (m === 0) continue;
      if (m === 4) { ap[d][q] = 0; continue; }
      let n = 0;
      for (const a of SUB) for (const b of SUB) {
        const x = d === 0 ? X.f[i + 1] : X.f[i] + a * X.d[i];
        const y = d === 1 ? Y.f[j + 1] : Y.f[j] + (d === 0 ? a : b) * Y.d[j];
        const z = d === 2 ? Z.f[k + 1] : Z.f[k] + b * Z.d[k];
        if (!solidAt(x, y, z)) n++;
      }
      ap[d][q] = n / 9;
    }
  };
  for (let k = k0; k < k1; k++) cellPlane(k);
  return { phi, ap };
}

/* Solid flag, wall-face cleanup, forced-open fixed-cell faces and counts
 * (src/air_cfd.js:114, 132-139). `fixed` holds [q, axis, side] records. */
function finishMasks(g, phi, ap, t, fixed) {
  const NX = g.NX, NY = g.NY, NZ = g.NZ, N = g.N, NXY = NX * NY;
  const a0 = ap[0], a1 = ap[1], a2 = ap[2], a3 = ap[3];
  // one plane per call (compiled after the first few); the + face of axis d
  // closes when either cell is solid, the z one also drops its baffle slope,
  // as does a z face on the top boundary or next to a fixed-state cell
  const plane = k => {
    const kin = k + 1 < NZ;
    for (let j = 0; j < NY; j++) {
      const jin = j + 1 < NY;
      for (let i = 0; i < NX; i++) {
        const q = (k * NY + j) * NX + i, s1 = t[q] === 1;
        if (i + 1 < NX && (s1 || t[q + 1] === 1)) a0[q] = 0;
        if (jin && (s1 || t[q + NX] === 1)) a1[q] = 0;
        if (!kin) a3[q] = 0;
        else if (s1 || t[q + NXY] === 1) { a2[q] = 0; a3[q] = 0; }
        if (t[q] >= 3 || (kin && t[q + NXY] >= 3)) a3[q] = 0;
      }
    }
  };
  for (let k = 0; k < NZ; k++) plane(k);
  for (const [q, d, s] of fixed) {
    if (s > 0) { const q2 = q + (d === 0 ? 1 : d === 1 ? NX : NXY); if (t[q2] === 0) ap[d][q] = 1; }
    else if (t[q - 1] === 0) ap[0][q - 1] = 1;
  }
  let nSolid = 0, nCut = 0;
  const count = (qa, qb) => {
    for (let q = qa; q < qb; q++) {
      if (t[q] === 1) nSolid++;
      else if (t[q] === 0 && (phi[q] < 1 || a0[q] < 1 || a1[q] < 1 || a2[q] < 1)) nCut++;
    }
  };
  for (let k = 0; k < NZ; k++) count(k * NXY, (k + 1) * NXY);
  return { nSolid, nCut };
}

/* Cut-cell masks of a plain body (no baffles or engine cells), e.g. the sphere. */
function cfdGeomBody(g, body) {
  const { phi, ap } = classifyCells(g, (x, y, z) => body.solidAt(x, Math.abs(y), z));
  const t = new Uint8Array(g.N);
  for (let q = 0; q < g.N; q++) t[q] = phi[q] < 0.2 ? 1 : 0;
  const c = finishMasks(g, phi, ap, t, []);
  return { t, phi, ap, nSolid: c.nSolid, nCut: c.nCut, volW: false, volH: false };
}

/* Solid test of an aircraft at a point (src/air_cfd.js:50-66): nacelle
 * envelopes (max meridian radius at 301 stations), fuselage, and the wing,
 * deflected tail and fin where they are at least 1.5 cells thick. */
function aircraftSolid(ac, G, g, deltaDeg) {
  // nacelle envelope: max meridian radius sampled at 301 stations
  const envs = [];
  for (const e of ac.engines) {
    if (e.internal || !e.meridian) continue;
    const mer = e.meridian, xmax = mer[mer.length - 1][0], n = 300, R = new Float32Array(n + 1);
    for (let k = 0; k <= n; k++) {
      const xe = xmax * k / n;
      let r = 0;
      for (let s = 0; s < mer.length - 1; s++) {
        const xa = mer[s][0], ra = mer[s][1], xb = mer[s + 1][0], rb = mer[s + 1][1];
        if (xe >= Math.min(xa, xb) - 1e-9 && xe <= Math.max(xa, xb) + 1e-9)
          r = Math.max(r, xb > xa ? ra + (rb - ra) * (xe - xa) / (xb - xa) : Math.max(ra, rb));
      }
      R[k] = r;
    }
    envs.push({ e, xmax, n, R });
  }
  const ht = ac.htail, dR = (deltaDeg || 0) * DEG;
  const hinge = [ht.xLE(0) + 0.25 * ht.cr, ht.o.z0], cR = Math.cos(dR), sR = Math.sin(dR);
  // a lifting surface thinner than 1.5 cells on average is a baffle only
  const thick = Sx => 0.5 * (Sx.o.tcR * Sx.cr + Sx.o.tcT * Sx.ct) >= 1.5 * g.h;
  const volW = thick(ac.wing), volH = thick(ht), volV = thick(ac.vtail);
  const solidAt = (x, y, z) => {
    const ay = Math.abs(y);
    for (const v of envs) {
      const xe = x - v.e.x0;
      if (xe < 0 || xe > v.xmax) continue;
      const f = xe / v.xmax * v.n, k = Math.min(v.n - 1, Math.floor(f)), r = v.R[k] + (v.R[k + 1] - v.R[k]) * (f - k);
      if (Math.hypot(ay - v.e.y, z - v.e.z) < r) return true;
    }
    if (G.inFus(x, ay, z)) return true;
    if (volW && ay >= G.sW * 0.9 && G.inLift(ac.wing, G.sW, G.bW, x, ay, z)) return true;
    if (volH && ay >= G.sH * 0.9) {
      const dx = x - hinge[0], dz = z - hinge[1];
      if (G.inLift(ht, G.sH, G.bH, hinge[0] + dx * cR - dz * sR, ay, hinge[1] + dx * sR + dz * cR)) return true;
    }
    return volV && G.inVt(x, ay, z);
  };
  return { solidAt, volW, volH, ht, dR, hinge };
}

/* Cut-cell masks of an aircraft (src/air_cfd.js:50-141): nacelle envelopes,
 * fuselage, thick wing / tail / fin, thin-surface baffles on the mean camber
 * surface with their slope dz/dx in ap[3], fixed-state inlet (t=3) and nozzle
 * (t=4 bypass, t=5 core) cells, wall-face cleanup. `pre` ({phi, ap}) takes the
 * classification of every plane from elsewhere (the parallel build). */
function cfdGeom(ac, G, g, deltaDeg, pre) {
  const NX = g.NX, NY = g.NY, NZ = g.NZ, X = g.X, Y = g.Y, Z = g.Z, sp = ac.spec;
  const { solidAt, volW, volH, ht, dR, hinge } = aircraftSolid(ac, G, g, deltaDeg);
  const { phi, ap } = pre || classifyCells(g, solidAt);

  // thin-surface baffle: close the z face nearest the mean camber surface by
  // the planform coverage, store the camber slope (incl. incidence, deflection)
  const zMean = (S, sA, x, ay) => {
    if (ay > S.semi || ay < sA) return null;
    const c = S.chord(ay), xc = (x - S.xLE(ay)) / c;
    if (xc < 0 || xc > 1) return null;
    return S.o.z0 + (ay - sA) * Math.tan((S.o.dih || 0) * DEG) + nacaC(S.o.camber || 0, xc) * c - (xc - 0.25) * c * Math.tan((S.o.inc || 0) * DEG);
  };
  const baffle = (S, sA, rot) => {
    for (let j = 0; j < NY; j++) {
      if (Y.f[j] > S.semi || Y.f[j + 1] < sA) continue;
      baffleRow(S, sA, rot, j);
    }
  };
  const baffleRow = (S, sA, rot, j) => {
    {
      for (let i = 0; i < NX; i++) {
        let cov = 0, zs = 0, ss = 0;
        for (const a of SUB) for (const b of SUB) {
          const x = X.f[i] + a * X.d[i], y = Y.f[j] + b * Y.d[j];
          let zm = zMean(S, sA, x, y);
          if (zm === null) continue;
          const e = 0.02 * S.chord(Math.min(y, S.semi)), zp = zMean(S, sA, x + e, y), zn = zMean(S, sA, x - e, y);
          let sl;
          if (zp !== null && zn !== null) sl = (zp - zn) / (2 * e);
          else if (zp !== null) sl = (zp - zm) / e;
          else if (zn !== null) sl = (zm - zn) / e;
          else sl = 0;
          if (rot) { zm -= (x - hinge[0]) * Math.tan(dR); sl -= Math.tan(dR); }
          cov++; zs += zm; ss += sl;
        }
        if (!cov) continue;
        const zm = zs / cov;
        let k = 0;
        while (k < NZ - 2 && Z.c[k + 1] <= zm) k++;
        const q = (k * NY + j) * NX + i;
        ap[2][q] = Math.min(ap[2][q], 1 - cov / 9);
        ap[3][q] = ss / cov;
      }
    }
  };
  baffle(ac.wing, G.sW, false);
  baffle(ht, G.sH, true);

  const t = new Uint8Array(g.N);
  for (let q = 0; q < g.N; q++) t[q] = phi[q] < 0.2 ? 1 : 0;
  // fixed-state cells: one layer upstream of each nozzle exit plane, one layer
  // downstream of the nacelle inlet face; a padded ring if the plane is missed
  const fixed = [];
  const layerAt = xv => { let i = 0; while (i < NX - 1 && X.f[i + 1] <= xv) i++; return i; };
  for (const e of ac.engines) {
    for (const ex of sp.exitsGeom) {
      const ty = ex.key === 'bypass' ? 4 : 5, i = layerAt(e.x0 + ex.x - 1e-3);
      const mark = pad => {
        let n = 0;
        const plane = k => {
          for (let j = 0; j < NY; j++) {
            const r = Math.hypot(Y.c[j] - e.y, Z.c[k] - e.z);
            if (r <= ex.R + pad && r >= (ex.Rin || 0) - pad) { const q = (k * NY + j) * NX + i; t[q] = ty; phi[q] = 1; fixed.push([q, 0, 1]); n++; }
          }
        };
        for (let k = 0; k < NZ; k++) plane(k);
        return n;
      };
      if (!mark(0)) mark(0.55 * Math.max(Y.d[0], Z.d[0], g.h));
    }
    if (!e.internal) {
      const i = layerAt(e.x0 + sp.lip.x + 1e-3);
      const mark = pad => {
        let n = 0;
        const plane = k => {
          for (let j = 0; j < NY; j++) {
            const r = Math.hypot(Y.c[j] - e.y, Z.c[k] - e.z);
            if (r <= 0.9 * sp.lip.r + pad) { const q = (k * NY + j) * NX + i; t[q] = 3; phi[q] = 1; fixed.push([q, 0, -1]); n++; }
          }
        };
        for (let k = 0; k < NZ; k++) plane(k);
        return n;
      };
      if (!mark(0)) mark(0.55 * g.h);
    }
  }
  const c = finishMasks(g, phi, ap, t, fixed);
  return { t, phi, ap, nSolid: c.nSolid, nCut: c.nCut, volW, volH };
}

/* Force stencil: fluid cells whose aperture difference across x or z is
 * non-zero (src/air_cfd.js:143-152). Returned as parallel typed arrays (the
 * form the device reduction takes) plus the reference's object list. */
function cfdForceList(g, geo) {
  const L = [], NX = g.NX, NXY = NX * g.NY;
  for (let k = 0; k < g.NZ; k++) for (let j = 0; j < g.NY; j++) for (let i = 0; i < g.NX; i++) {
    const q = (k * g.NY + j) * NX + i;
    if (geo.t[q] !== 0) continue;
    const axm = i > 0 ? geo.ap[0][q - 1] : 1, azm = k > 0 ? geo.ap[2][q - NXY] : 1;
    const cx = (geo.ap[0][q] - axm) * g.Y.d[j] * g.Z.d[k], cz = (geo.ap[2][q] - azm) * g.X.d[i] * g.Y.d[j];
    if (cx !== 0 || cz !== 0) L.push({ q, cx, cz, x: g.X.c[i], z: g.Z.c[k] });
  }
  return L;
}
function forceListArrays(list) {
  const n = list.length, q = new Int32Array(n), cx = new Float64Array(n), cz = new Float64Array(n), x = new Float64Array(n), z = new Float64Array(n);
  for (let s = 0; s < n; s++) { const f = list[s]; q[s] = f.q; cx[s] = f.cx; cz[s] = f.cz; x[s] = f.x; z[s] = f.z; }
  return { n, q, cx, cz, x, z };
}

/* Trefftz plane of an aircraft: 0.12 L behind the later of the fuselage end
 * and the tail trailing edge (src/air_cfd.js:561). */
function trefftzX(ac) { return Math.max(ac.fus.L, ac.htail.xLE(ac.htail.semi) + ac.htail.ct) + 0.12 * ac.fus.L; }

/* Freestream conserved state at Mach M and incidence alpha (src/air_cfd.js:523-524). */
function freestream(M, alphaDeg) {
  const a = alphaDeg * DEG;
  return [1, M * Math.cos(a), 0, M * Math.sin(a), 1 / (CG * (CG - 1)) + 0.5 * M * M];
}

module.exports = {
  CG, CFD_GRIDS, SUB, cfdAxis, axisFromFaces, gridFromFaces, cfdGrid, cidx, classifyCells, finishMasks,
  cfdGeomBody, aircraftSolid, cfdGeom, cfdForceList, forceListArrays, trefftzX, freestream,
};
