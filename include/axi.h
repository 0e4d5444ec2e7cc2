/* ============================================================================
 * axi.h — C ABI of the MI355X axisymmetric engine solver (part of libaero3d.so).
 *
 * Replaces the reference's GPUSolver / CPUSolver (src/cfd_gpu.js:207-323,
 * src/cfd_cpu.js:50-214): 5 conserved variables [rho, rho u, rho v, E, rho Y]
 * on a W x H (x, r) cut-cell grid, variable gamma = 1.4 - 0.07 Y, HLLC +
 * MUSCL-minmod, Riemann-invariant farfield, mixing-length r fluxes, blade-row
 * body forces, combustor heat source, SSP-RK2 with local dt. FP32 state on the
 * device. Conventions as aero3d.h (0 = ok, < 0 = error, axi_last_error()).
 * ========================================================================== */
#ifndef AXI_H
#define AXI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct axi_solver axi_solver;

/* Grid of buildCFDGrid (src/cfd_grid.js:28-43): W x H cells, Float64 faces;
 * W, H >= 2 and W*H <= 2^28. */
typedef struct {
  int W, H;
  const double *xf, *rf;
} axi_grid;

/* Zones of buildZones (src/cfd_grid.js:81-191), N = W*H entries each (row j, column i at j*W+i). */
typedef struct {
  const uint8_t *solid, *comp;
  const float *phi, *ae, *an, *wgt, *lmix, *tgt, *rate, *yrate, *mfd;
} axi_zones;

/* new GPUSolver(grid, zones, comps, fs, Q0, opt) (src/cfd_gpu.js:207): comps =
 * 32 doubles (compParams, src/cfd_grid.js:260-276), fs = {r, u, p} freestream,
 * Q0 = 5 planes of N doubles. */
int axi_create(const axi_grid *g, const axi_zones *z, const double comps[32], const double fs[3], const double *Q0,
               axi_solver **out);
int axi_destroy(axi_solver *s);
const char *axi_last_error(const axi_solver *s);
/* opt (src/cfd_cpu.js:53, 63, 198-202): cfl, order (1|2), visc (0|1), hf heating
 * value (<= 0 keeps 0.99 * 43e6), global (0|1) and dtRatio (0 = off) dt caps. */
int axi_set_opt(axi_solver *s, double cfl, int order, int visc, double hf, int global, double dtRatio);
/* Kernel form of the steps without a dt cap (no reference counterpart, a
 * tuning knob): 0 a face kernel and a cell kernel per stage with the fluxes
 * in memory; 1 .. 4 one fused kernel per stage, tiles of 63 x (3, 7, 9, 5)
 * cells in blocks of (4, 8, 16, 4) waves with the primitives and face fluxes
 * in LDS. The same arithmetic, the same bits. Default at create: 3 when its
 * tiles make between half and all of one block per compute unit, else 0. */
int axi_set_fused(axi_solver *s, int form);
/* Times every kernel form over `iters` iterations, keeps the fastest and
 * restores the state, diagnostics and iteration count; outputs may be null. */
int axi_autotune(axi_solver *s, int iters, int *form, float *ms_best);
/* setComps / setZones / setFs (src/cfd_gpu.js:252-265). */
int axi_set_comps(axi_solver *s, const double comps[32]);
int axi_set_zones(axi_solver *s, const axi_zones *z);
int axi_set_fs(axi_solver *s, const double fs[3]);
/* setState(Q) (src/cfd_gpu.js:266-274): 5 planes of N doubles. */
int axi_set_state(axi_solver *s, const double *Q);
/* step(n) (src/cfd_cpu.js:193-210), enqueued on the solver's stream. */
int axi_step(axi_solver *s, int n);
int axi_iter(const axi_solver *s, int *it);
/* read() (src/cfd_gpu.js:303-316): Q as 5 planes of N floats, and the
 * diagnostics src (blade-row energy or fuel source per fluid volume) and mfx
 * (mean axial face mass flux) of the last stage. Any pointer may be null. Waits. */
int axi_read(axi_solver *s, float *Q, float *src, float *mfx, float *dt);
int axi_sync(axi_solver *s);
/* n iterations bracketed by HIP events on the solver's stream. */
int axi_time_steps(axi_solver *s, int n, float *ms);

/* Readback monitors of the application's convergence loop on the device, in
 * FP64 (src/app_cfd.js:361-404): what the loop otherwise computes on the host
 * from a full read() after each chunk of iterations.
 *   F, Fb, Fw, mfuel, pw  cfdForces with the solver's own sources
 *                         (src/cfd_post.js:44-75; F = Fb + Fw)
 *   bad                   non-physical cells of the divergence check
 *   res_sum2, res_n       sum of ((rho - rho_prev) / rho_prev)^2 and the count
 *                         over every second non-solid cell, against the state
 *                         of the previous axi_monitor call (src/app_cfd.js:368-371);
 *                         res_valid = 0 on the first call and after setState
 *   sec_*                 per section of axi_monitor_sections: mass flow, sum
 *                         of max(mf, 0) and of max(mf, 0) Tt (cfdSection with
 *                         the solver's mass flux, src/cfd_post.js:83-102)
 * Per-cell terms are the host's double operations; the grid sums differ from
 * the host's sequential ones only in their order. Waits. */
typedef struct {
  double F, Fb, Fw, mfuel, pw[8];
  double res_sum2, res_n;
  int res_valid, nsec;
  double sec_mdot[4], sec_w[4], sec_wTt[4];
  int bad;                        /* cells k = 0, 7, 14, ... with rho or E not positive (the
                                     divergence check of src/app_cfd.js:377); > 0: diverged */
} axi_monitor_out;
int axi_monitor(axi_solver *s, const double comps[32], double P0, axi_monitor_out *out);
/* Sections for axi_monitor (at most 4): column col[q] and row mask rows[q*H + j]
 * (1 where cfdSection counts row j: non-solid and in the section's stream). */
int axi_monitor_sections(axi_solver *s, int nsec, const int *col, const uint8_t *rows);

/* Section sums of the current state on the device, for cfdSection
 * (src/cfd_post.js:83-102; the control loop's section averages) and the
 * station averages (tools/stations.js:8-16): per section q, column col[q]
 * and rows[q*H + j] = 1 where row j counts (non-solid, in the section's
 * stream). The cell mass flow is the solver's mean axial face mass flux times
 * the annulus area. sums[11 q + c], FP64 in the host's row order:
 *   0 mass flow, 1 w = sum of max(mf, 0), 2 area (aperture-weighted),
 *   3 sum p dA, 4..10 the w-weighted sums of Tt, Pt, T, p, |V|, M, Y.
 * cfdSection's averages are the weighted sums over (w or 1), its P without
 * positive flow pA / (A or 1); a station's Tt, Pt the weighted sums over w.
 * nsec <= 256. Waits. */
int axi_sections(axi_solver *s, int nsec, const int *col, const uint8_t *rows, double *sums);
/* Display fields of the current state on the device (cfdDerived,
 * src/cfd_post.js:10-40). axi_set_display_zones gives the zone codes they
 * need once (stream: Int8, -1 outside the engine; mat: material of solid
 * cells). disp = 1: cells under half full count as solid. fields: 13 planes
 * of N floats, in the order T, Tt, P, Pt, V, M, rho, Vx, Vth, Vr, Y, vort, ds
 * (NaN on solid cells, Vth 0); reg: 1 solid, 0 outside, 2 + stream inside;
 * mat: the material code on solid cells, 0 elsewhere. Any output may be
 * null. Every value is the reference's double arithmetic rounded to FP32
 * once (Pt and ds through pow / log, within an ulp of a JS engine's). Waits. */
int axi_set_display_zones(axi_solver *s, const int8_t *stream, const uint8_t *mat);
int axi_derived(axi_solver *s, int disp, double T0, double P0, float *fields, uint8_t *reg, uint8_t *mat);
/* Grid sequencing (interpState, src/app_cfd.js:321-358): the state of src
 * carried onto dst's grid on the device and made dst's current state
 * (bilinear in the cell centres over src's non-solid cells; where none has
 * weight, the nearest non-solid cell of a 7 x 7 window; where there is none,
 * P0 / (287.05 T0) density and P0 / 0.4 energy at rest). FP64 arithmetic,
 * rounded to FP32 once. Both solvers on one device; waits for src's stream. */
int axi_transfer(axi_solver *dst, const axi_solver *src, double T0, double P0);

/* ---------------------------------------------------------------------------
 * Engine field model (FieldModel, src/field.js:255-488) and the solver's
 * initial state (initialState, src/cfd_grid.js:279-305) on the device.
 *
 * The model is given as its tables: what the FieldModel constructor derives
 * from the engine cycle (stream knots after assignMach, the processed exits,
 * the inlet face velocity, the hot gas) and the engine geometry it samples
 * (stream walls, body polygons, combustor, mixer, afterburner, lip and outer
 * line). One flat array of doubles: a header of AXI_FM_HEAD entries at the
 * indices below, then a data area that the header's offsets index (counts
 * and offsets are integer-valued doubles, undefined cycle values are NaN).
 * -------------------------------------------------------------------------*/
enum {
  AXI_FM_VERSION = 0,   /* 1 */
  AXI_FM_TYPE = 1,      /* 0 turbojet, 1 high-bypass turbofan, 2 low-bypass turbofan */
  AXI_FM_S = 2,         /* engine scale s */
  AXI_FM_NSTREAM = 3,   /* streams in the model's classification order, 1..4 */
  AXI_FM_STREAM = 4,    /* 4 x 10: name (0 inlet, 1 bypass, 2 core, 3 ab), x0, x1, lower wall (offset,
                           points), upper wall (offset, points), knots (offset, count), 0; a wall is
                           (x, r) pairs, a knot 10 doubles: x, Tt, Pt, M, swirl deg, g, cp, R, mdot, noz */
  AXI_FM_NBODY = 44,    /* solid polygons, 0..16 */
  AXI_FM_BODY = 45,     /* 16 x 6: x min, x max, r min, r max, polygon (offset, points) */
  AXI_FM_COMB = 141,    /* stream index, x0, xd, x1, turbEnd, n, cannular, inner liner (offset, points),
                           outer liner (offset, points) */
  AXI_FM_MIXER = 152,   /* present, x, coreFrac, lobes */
  AXI_FM_AB = 156,      /* present, stream index, xFH, x1, rings (offset, count) */
  AXI_FM_CYCLE = 162,   /* Tt3, Tt4, Tt5, Tt6A, Tt7, Tt13, abOn */
  AXI_FM_FL = 169,      /* freestream V0, T0, P0, Tt0, Pt0 */
  AXI_FM_LIP = 174,     /* lip x, r, rOut, outer line (offset, points), outerEnd */
  AXI_FM_COLD = 180,    /* cold gas g, cp */
  AXI_FM_NEXIT = 182,   /* exits, 0..4, in the model's order (largest radius first) */
  AXI_FM_EXIT = 183,    /* 4 x 12: x, R, spread, Qm, Qh, under, Ls, P, Pt, g, cp, dTt */
  AXI_FM_HOT = 231,     /* Uface, maxdT, hot gas present, hot gas g, cp */
  AXI_FM_HEAD = 240
};
typedef struct axi_field axi_field;
/* new FieldModel(...): the table is checked (version, counts, every offset and
 * extent inside the array) and kept on the device. */
int axi_field_create(const double *tab, int ntab, axi_field **out);
int axi_field_destroy(axi_field *f);
const char *axi_field_last_error(const axi_field *f);
/* FieldModel.prototype.sample (src/field.js:318-488) at n points (x, r, th):
 * out[14 p + c] = region (0 outside the engine, 1 solid, 2 + stream index),
 * then T, P, Tt, Pt, V, Vx, Vth, M, a, rho, g, cp, R (NaN for a solid point).
 * FP64, the reference's operations in its order; pow / exp / sin / cos / tan /
 * log within an ulp of a JS engine's. Waits. */
int axi_field_sample(axi_field *f, const double *x, const double *r, const double *th, int n, double *out);
/* initialState(grid, zones, spec, model, fl, uniform) (src/cfd_grid.js:279-305)
 * on the grid with faces xf (W + 1) and rf (H + 1) and the zones' solid (u8)
 * and stream (i8, index into the model's stream order, -1 outside): Q = 5
 * planes of N = W*H doubles [rho, rho u, 0, E, rho Y]. W, H >= 1 and
 * W*H <= 2^28. Waits. */
int axi_initial_state(axi_field *f, int W, int H, const double *xf, const double *rf, const uint8_t *solid,
                      const int8_t *stream, int uniform, double *Q);

#ifdef __cplusplus
}
#endif
#endif
