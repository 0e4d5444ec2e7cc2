'use strict';
/* ============================================================================
 * cfd_post.js — host side of the axisymmetric post-processing (drop-in for
 * the control part of src/cfd_post.js:104-126 and for the station averages of
 * tools/stations.js:8-16). The sums over the state run on the device
 * (AxiHIP.monitor: thrust, blade-row powers, residual, control sections;
 * AxiHIP.sections: any section or station; AxiHIP.derived: display fields).
 * This module holds what is left on the host: where a section is (column and
 * rows), averages from the device sums, the closed-loop exit-temperature
 * control and the spool power ratios.
 * ========================================================================== */
const clamp = (v, a, b) => (v < a ? a : v > b ? b : v);           // src/field.js:4

/* Column whose centre is nearest to x (the right one of two at equal distance). */
function nearestColumn(xc, x) {
  let best = 0;
  for (let i = 1; i < xc.length; i++) if (Math.abs(xc[i] - x) <= Math.abs(xc[best] - x)) best = i;
  return best;
}

/* A section of cfdSection (src/cfd_post.js:83-102) for the device sums: the
 * column at x and the rows it counts (non-solid; in stream si unless si is null). */
function sectionRows(grid, zones, x, si) {
  const { W, H } = grid, col = nearestColumn(grid.xc, x), rows = new Uint8Array(H);
  for (let j = 0; j < H; j++) {
    const k = j * W + col;
    rows[j] = !zones.solid[k] && (si === null || zones.stream[k] === si) ? 1 : 0;
  }
  return { col, rows };
}

/* The section record of cfdSection from AxiHIP.sections' sums s (mass flow,
 * w, area, sum p dA, w-weighted Tt, Pt, T, p, |V|, M, Y): mass-weighted
 * averages over the positive mass flow (or over 1 when there is none, with
 * the area-weighted pressure then). */
function sectionAverages(s, x) {
  const [mdot, w, A, pA, Tt, Pt, T, P, V, M, Y] = s, d = w || 1;
  return { x, mdot, A, Tt: Tt / d, Pt: Pt / d, T: T / d, P: w ? P / d : pA / (A || 1), V: V / d, M: M / d, Y: Y / d };
}

/* {mdot, Tt} of a control section from the device monitor's sums {mdot, w, wTt}. */
const sectionFromSums = s => ({ mdot: s.mdot, Tt: s.wTt / (s.w || 1) });

const SPOOLS = { hbtf: [[[1, 2], 6], [[3], 5]], tj: [[[3], 5]], lbtf: [[[1], 6], [[3], 5]] };   // src/cfd_post.js:106
/* Turbine-to-compressor power ratio per spool (src/cfd_post.js:124-126). */
function spoolRatios(type, pw, etam) {
  return SPOOLS[type].map(([cs, t]) => { const Pc = cs.reduce((a, id) => a + pw[id], 0), Pt = -pw[t]; return Pc > 0 && Pt > 0 ? etam * Pt / Pc : NaN; });
}

/* The control stations of cfdControl: heat source id, section x and stream,
 * target exit temperature (combustor; afterburner when lit). */
function controlStations(ctl) {
  const out = [{ id: 4, x: ctl.comb.x1 - 0.004 * ctl.s, si: ctl.order.indexOf(ctl.comb.stream), target: ctl.Tt4 }];
  if (ctl.ab && ctl.abOn) out.push({ id: 7, x: ctl.ab.x1 - 0.004 * ctl.s, si: ctl.order.indexOf(ctl.ab.stream), target: ctl.Tt7 });
  return out;
}

/* cfdControl's gain update (src/cfd_post.js:110-123) with the section of each
 * control station from secOf(st) ({mdot, Tt}): the gain K of comps[id*4+1]
 * moves toward the target exit temperature, by at most 3 % per 100
 * iterations, within [0.5, 2.5]. ctl = {s, comb: {x1, stream}, ab, order,
 * Tt4, Tt7, abOn}: the spec / operating-point fields the reference reads. */
function controlFromSections(type, ctl, secOf, comps, pw, etam, dIt) {
  const lim = 0.03 * Math.min(1, (dIt || 100) / 100);
  const out = { ratios: spoolRatios(type, pw, etam), Tt4: NaN, Tt7: NaN };
  for (const st of controlStations(ctl)) {
    const sec = secOf(st), id = st.id, target = st.target;
    const T0 = comps[id * 4], K = comps[id * 4 + 1] || 1;
    if (sec.mdot > 0 && sec.Tt > T0 + 20 && target > T0 + 20) comps[id * 4 + 1] = clamp(K * clamp(Math.pow((target - T0) / (sec.Tt - T0), 0.3), 1 - lim, 1 + lim), 0.5, 2.5);
    out[id === 4 ? 'Tt4' : 'Tt7'] = sec.Tt;
  }
  return out;
}

/* Station table of the long-run report (tools/stations.js:8-16): id, x, stream. */
const STATIONS = {
  hbtf: [['2', 0.58, 'inlet'], ['13', 0.95, 'bypass'], ['25', 1.47, 'core'], ['3', 2.05, 'core'], ['4', 2.43, 'core'], ['45', 2.62, 'core'], ['5', 3.08, 'core'], ['19', 2.19, 'bypass']],
  tj: [['2', 0.36, 'core'], ['3', 1.55, 'core'], ['4', 2.21, 'core'], ['5', 2.62, 'core'], ['7', 4.29, 'core'], ['9', 4.89, 'core']],
  lbtf: [['2', 0.3, 'inlet'], ['13', 0.9, 'bypass'], ['3', 1.55, 'core'], ['4', 2.0, 'core'], ['5', 2.55, 'core'], ['6A', 3.0, 'ab'], ['9', 4.94, 'ab']],
};

/* Station averages on the device: at each station's column, over its
 * stream's rows, the mass flow and the Tt, Pt weighted by the positive cell
 * mass flow (from the mean axial face mass flux). S: an AxiHIP; list: an
 * engine type ('tj', 'hbtf', 'lbtf') or [[id, x, stream], ...]. */
function stations(S, grid, zones, order, list) {
  const L = typeof list === 'string' ? STATIONS[list] : (list || STATIONS.tj);
  const secs = L.map(([, x, sn]) => sectionRows(grid, zones, x, order.indexOf(sn)));
  return S.sections(secs).map((s, q) => ({ id: L[q][0], x: L[q][1], mdot: s[0], Tt: s[4] / s[1], Pt: s[5] / s[1] }));
}

module.exports = { clamp, nearestColumn, sectionRows, sectionAverages, sectionFromSums, spoolRatios, controlStations, controlFromSections,
  STATIONS, stations };
