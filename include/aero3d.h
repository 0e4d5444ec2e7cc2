/* ============================================================================
 * aero3d.h — C ABI of the MI355X 3D Euler cut-cell solver (libaero3d.so).
 *
 * One handle = one solve: the library owns the device copies of the grid,
 * the cut-cell masks and three FP32 structure-of-arrays state buffers, plus a
 * HIP stream on which every call of the handle is ordered. Host arrays stay
 * owned by the caller and are copied (read() copies out, unlike the live
 * Float32Array references of the reference's CfdCPU.read).
 *
 * Conventions: every function returns 0 on success or a negative code
 * (AERO3D_E_*, or -(1000 + hipError_t) for a HIP failure); the message of the
 * last failure is returned by aero3d_last_error(). A handle must not be used
 * by two threads at once. Entry points are listed with the reference
 * interface each one replaces (file:line in the reference checkout).
 * ========================================================================== */
#ifndef AERO3D_H
#define AERO3D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AERO3D_OK 0
#define AERO3D_E_ARG (-1)        /* bad argument (null pointer, size, option)   */
#define AERO3D_E_NODEV (-2)      /* no usable HIP device                        */
#define AERO3D_E_STATE (-3)      /* call out of order (e.g. no force list set)  */
#define AERO3D_E_HIP_BASE (-1000)

typedef struct aero3d_solver aero3d_solver;

/* Grid of the half domain y >= 0: face coordinates of each axis (Float64,
 * NX+1 / NY+1 / NZ+1 entries), as the reference's g.X.f / g.Y.f / g.Z.f
 * (src/air_cfd.js:36). Cell widths and centres are rounded to FP32 exactly as
 * the reference grid stores them. NX, NZ >= 1, NY >= 2 (the mirror ghosts of
 * rows -1, -2 are rows 0, 1) and NX*NY*NZ <= 2^30; other sizes are rejected
 * with AERO3D_E_ARG. */
typedef struct {
  int NX, NY, NZ;
  const double *fx, *fy, *fz;
} aero3d_grid;

/* new CfdCPU(g, geo) / new CfdGPU(g, geo) (src/air_cfd.js:157, 367).
 * t: cell flags (0 fluid, 1 solid, 3 inlet sink, 4 bypass jet, 5 core jet),
 * phi: fluid volume fraction, ap0..ap2: + face apertures, ap3: baffle slope of
 * the + z face; all N = NX*NY*NZ entries in cidx order (x fastest,
 * src/air_cfd.js:45). Throws (returns < 0) when no device is usable. */
int aero3d_create(const aero3d_grid *g, const uint8_t *t, const float *phi, const float *ap0, const float *ap1,
                  const float *ap2, const float *ap3, aero3d_solver **out);
/* CfdGPU.dispose (src/air_cfd.js:438). Null is accepted. */
int aero3d_destroy(aero3d_solver *s);
/* Message of the last failure on this handle (or of the last failed create
 * when s is null). */
const char *aero3d_last_error(const aero3d_solver *s);

/* setFree(F) (src/air_cfd.js:166): freestream conserved state [rho, rho u, rho v, rho w, E]. */
int aero3d_set_free(aero3d_solver *s, const double F[5]);
/* setFix(fix) (src/air_cfd.js:167): fixed states of cell flags 3, 4, 5, row-major 3 x 5. */
int aero3d_set_fix(aero3d_solver *s, const double fix[15]);
/* opt = {cfl, order} (src/air_cfd.js:162), read at every step. order is 1 or 2. */
int aero3d_set_opt(aero3d_solver *s, double cfl, int order);
/* reset() (src/air_cfd.js:168): state := F everywhere, it := 0. */
int aero3d_reset(aero3d_solver *s);
/* step(n) (src/air_cfd.js:252-258, 414-428): n SSP-RK2 iterations, enqueued
 * on the handle's stream; returns without waiting. */
int aero3d_step(aero3d_solver *s, int n);
/* Iteration count `it`. */
int aero3d_iter(const aero3d_solver *s, int *it);
/* read() (src/air_cfd.js:259, 429-437): copy the current state (and the
 * local dt of the last iteration when dt is non-null) to caller arrays of N
 * floats; waits for the stream. Any output pointer may be null. */
int aero3d_read(aero3d_solver *s, float *r, float *ru, float *rv, float *rw, float *E, float *dt);
/* GPUSolver.setState (src/cfd_gpu.js:259-261) for the 3D solver: overwrite
 * the current state from caller arrays of N floats. */
int aero3d_write(aero3d_solver *s, const float *r, const float *ru, const float *rv, const float *rw, const float *E);
/* Wait for all work of the handle. */
int aero3d_sync(aero3d_solver *s);

/* cfdForceList (src/air_cfd.js:143-152) uploaded once: n entries of cell
 * index q, aperture-jump areas cx, cz and cell centre x, z. */
int aero3d_set_force_list(aero3d_solver *s, const int32_t *q, const double *cx, const double *cz, const double *x,
                          const double *z, int n);
/* cfdForces(g, list, read(), alpha, M, S, xcg, zcg, cref) (src/air_cfd.js:443-452)
 * on the device state, FP64 sums: out = {CL, CD, Cm}. Waits for the stream. */
int aero3d_forces(aero3d_solver *s, double alphaDeg, double M, double S, double xcg, double zcg, double cref,
                  double out[3]);
/* Trefftz-plane induced drag of CFD3.readback (src/air_cfd.js:560-570) at
 * plane xT, FP64 sums. Waits for the stream. */
int aero3d_trefftz(aero3d_solver *s, double xT, double alphaDeg, double M, double S, double *cdi);
/* Density residual of CFD3.readback (src/air_cfd.js:553-554): RMS over every
 * third fluid cell of rho minus the density snapshot, divided by the
 * iterations since the snapshot; then the snapshot is retaken. The first call
 * after create/reset only takes the snapshot and returns *res = NaN. */
int aero3d_residual(aero3d_solver *s, double *res);

/* The readback of the reference's controller, CFD3.readback
 * (src/air_cfd.js:550-571), on the device in one call: the density residual
 * since the previous readback / residual call (NaN at the first; the snapshot
 * is retaken), cfdForces with the force list (NaN without one), the Trefftz
 * induced drag at plane xT (NaN when xT is NaN) and the largest Mach number
 * over every second fluid cell. M is floored at 0.05 as the reference does.
 * FP64 sums with the reference's per-term arithmetic; waits. */
typedef struct {
  double alpha, M, S, xcg, zcg, cref;   /* cfdForces arguments (ac.ref), alpha in degrees */
  double xT;                            /* Trefftz plane x, NaN: none                      */
} aero3d_readback_in;
typedef struct {
  int it;                               /* iteration count at the readback                 */
  double res, CL, CD, Cm, CDi, Mmax;
} aero3d_readback_out;
int aero3d_readback(aero3d_solver *s, const aero3d_readback_in *in, aero3d_readback_out *out);

/* Display sampling of the current state at n points, on the device: the
 * reference's cfdStencil + cfdSample + cfdState (src/air_cfd.js:453-490) and
 * cfdVelAt (:637-646), which its pages run on read() output.
 * xyz: 3n coordinates of the full model (y of either sign; v is mirrored
 * below the symmetry plane). Per point: the trilinear stencil over fluid and
 * fixed-state cells, renormalised, or the nearest fluid cell of the 4 x 4 x 4
 * block when it holds no weight (idx / w: 8 entries each, corner bit 0 = +x,
 * bit 1 = +y, bit 2 = +z; ok: found); the primitives rho, u, v, w, p of the
 * blended conserved state (W: 5, NaN where not found); the dimensional state
 * P, T, rho, V, M, Tt, Pt, Cp, Taw, Vx, Vy, Vz from the freestream record fs
 * (state: 12); the velocity in units of the freestream speed (divided by
 * max(M, 1e-3)) inside the box of cell centres, without the nearest-cell
 * fallback (vel: 3, NaN where vok is 0). Any output pointer may be NULL (not
 * copied back); state needs fs. Stencils, primitives and velocities equal
 * the reference's bit for bit (FP32 weights, FP64 sums in its order); the
 * dimensional state agrees to a few ulps (sqrt, pow). n <= 2^26; each axis
 * needs >= 2 cells. Waits. */
typedef struct { double T, a, P, rho, q; } aero3d_probe_fs;   /* the reference's fs record: freestream T, a, P, rho, q */
typedef struct {
  int32_t *idx;      /* 8n  */
  float *w;          /* 8n  */
  uint8_t *ok;       /* n   */
  double *W;         /* 5n  */
  double *state;     /* 12n */
  double *vel;       /* 3n  */
  uint8_t *vok;      /* n   */
} aero3d_probe_out;
int aero3d_probe(aero3d_solver *s, int n, const double *xyz, const aero3d_probe_fs *fs, double M, const aero3d_probe_out *out);

/* No reference counterpart: a variant of aero3d_readback for callers that
 * cannot allocate device memory (code under stream capture, serving loops
 * that own a fixed arena). It works on device arrays the caller owns (no
 * handle) and calls no hipMalloc / hipFree: its block partials and sums live
 * in a static device buffer of about 58 KB per GPU, held until the process
 * ends, whose size does not depend on the field's (every sum is strided over
 * at most 1024 blocks); fields of up to 2^30 cells and force lists of up to
 * 2^30 entries, larger ones are rejected with AERO3D_E_ARG. The buffer is
 * shared by every call on a GPU: calls that overlap on one GPU race. Results
 * equal aero3d_readback's on the same state bit for bit (same kernels).
 * Q: 5 planes of ns floats; rsnap: the density snapshot (N floats, every third
 * cell read and rewritten with the current density); nit: iterations since
 * the snapshot (0: snapshot only, res = NaN); t, dy, dz, fq..z: device arrays;
 * xc: the FP32 x cell centres on the HOST (Trefftz plane lookup), needed with
 * dy, dz only when in->xT is not NaN. `stream` is a hipStream_t (null: the
 * default stream); the call waits for it. out->it is -1. Errors: message in
 * aero3d_last_error(NULL). */
typedef struct {
  int NX, NY, NZ;
  uint64_t ns;
  const float *Q;
  float *rsnap;
  int nit;
  const uint8_t *t;
  const float *dy, *dz;
  const float *xc;
  const int32_t *fq;
  const double *cx, *cz, *x, *z;
  int nf;
} aero3d_field;
int aero3d_readback_static(const aero3d_field *f, const aero3d_readback_in *in, aero3d_readback_out *out, void *stream);

/* Benchmark helper: run n iterations on the handle's stream bracketed by HIP
 * events recorded on that stream; *ms = elapsed device time. */
int aero3d_time_steps(aero3d_solver *s, int n, float *ms);
/* Stage-kernel shape (the tuning knob of the stage kernel, no reference
 * counterpart: the reference's shader has one fragment per cell,
 * src/air_cfd.js:405-413). ty = 4 or 8: tiles of 64 x ty columns marching z;
 * ty = 32: tiles of 32 x 8 columns (a wavefront covers two 32-column rows),
 * for grids whose x extent fills 32-column tiles better (same states bit for
 * bit as the other tile shapes); zchunk > 0: fixed chunks of zchunk planes, one block each; zchunk <= 0:
 * balanced launch of max(1, -zchunk) x (resident blocks of the device), each
 * block marching an equal share of one tile's planes. ty = 0: one thread per
 * cell, each face flux evaluated by both of its cells as the reference does.
 * Default zchunk 0 and ty 32 where 32-column tiles cover NX with fewer padded
 * columns than 64-column ones (NX 72, 95, 160, ...), else ty 8. */
int aero3d_set_tile(aero3d_solver *s, int ty, int zchunk);
int aero3d_get_tile(const aero3d_solver *s, int *ty, int *zchunk);
/* Autotuner of the stage-kernel shape: times every candidate (ty in {4, 8, 32},
 * zchunk in {0, -2, -3, 2, 3, 4, 6, 8}) over `iters` iterations on this handle's grid, keeps
 * the fastest, and restores the state, the local dt that read() returns and
 * the iteration count. Outputs may be
 * null; *ms_best is the best time per iteration. */
int aero3d_autotune(aero3d_solver *s, int iters, int *ty, int *zchunk, float *ms_best);
/* Device-memory footprint of the handle in bytes. */
int aero3d_footprint(const aero3d_solver *s, int64_t *bytes);

/* ---------------------------------------------------------------------------
 * Batched form (no reference counterpart: the reference steps one solve per
 * WebGL context, src/air_cfd.js:414-428). Many independent solves of uneven
 * grid sizes share one device arena; every iteration is one persistent
 * launch per RK stage over all cases, replayed from a hipGraph.
 * ------------------------------------------------------------------------- */
typedef struct aero3d_batch aero3d_batch;
/* Arena of `arena_bytes` device bytes for geometries and case states, at most
 * `max_slots` resident cases. */
int aero3d_batch_create(int64_t arena_bytes, int max_slots, aero3d_batch **out);
int aero3d_batch_destroy(aero3d_batch *b);
const char *aero3d_batch_last_error(const aero3d_batch *b);
/* Register a grid + cut-cell masks + force list (the inputs of new CfdCPU(g,
 * geo) and cfdForceList, src/air_cfd.js:143-157), shared read-only by every
 * case that uses it. *gid identifies it. */
int aero3d_batch_add_geometry(aero3d_batch *b, const aero3d_grid *g, const uint8_t *t, const float *phi, const float *ap0,
                              const float *ap1, const float *ap2, const float *ap3, const int32_t *fq, const double *fcx,
                              const double *fcz, const double *fx, const double *fz, int nf, int *gid);
int aero3d_batch_remove_geometry(aero3d_batch *b, int gid);
/* Device bytes one case on geometry gid occupies (state, snapshot, partials). */
int aero3d_batch_case_bytes(const aero3d_batch *b, int gid, int64_t *bytes);
/* Add a case on geometry gid: setFree(F), setFix(fix), opt {cfl, order},
 * reset() (src/air_cfd.js:162-168). Fails with AERO3D_E_STATE when no slot or
 * arena room is left. */
int aero3d_batch_add(aero3d_batch *b, int gid, const double F[5], const double fix[15], double cfl, int order, int *slot);
int aero3d_batch_remove(aero3d_batch *b, int slot);
/* Stage-kernel z chunk of every case (tiles of 64 x 8 or 32 x 8 columns): zchunk > 0
 * planes per work item; zchunk 0 (the default) from the load, each time the
 * case table changes: the longest of 8, 6, 4 and 3 planes that gives at
 * least four work items per resident block (long chunks for throughput when
 * many cases are in flight, short ones for latency when few are). Negative
 * values are rejected with AERO3D_E_ARG. */
int aero3d_batch_set_zchunk(aero3d_batch *b, int zchunk);
/* Tile width of the stage kernel: 64 or 32 columns for every case, or 0 (the
 * default) per case: 32 when 32-column tiles cover the case's x extent with
 * fewer padded columns than 64-column ones, else 64; cases of each width are
 * launched apart, so when either width would have fewer than two work items
 * per resident block every case takes 64 columns (one launch). States are the
 * same bit for bit either way. Other values are rejected with AERO3D_E_ARG. */
int aero3d_batch_set_tile_width(aero3d_batch *b, int width);
/* Cases are stepped in groups of about `cells` cells (default: all cases in
 * one group): a group's state and masks stay in the Infinity Cache across its
 * stages and iterations before the next group runs. */
int aero3d_batch_set_group_cells(aero3d_batch *b, int64_t cells);
/* Launch form of the stage kernel: 0 (default) one block per work item (the
 * grid is the group's item count; a graph is captured again when it changes),
 * 1 the device's resident blocks walking all items (one persistent launch per
 * stage whatever the table). Results are the same bits either way. */
int aero3d_batch_set_persistent(aero3d_batch *b, int on);
/* n SSP-RK2 iterations of every resident case (src/air_cfd.js:252-258), enqueued. */
int aero3d_batch_step(aero3d_batch *b, int n);
int aero3d_batch_sync(aero3d_batch *b);
int aero3d_batch_iter(const aero3d_batch *b, int slot, int *it);
/* read() of one case (src/air_cfd.js:259); waits. */
int aero3d_batch_read(aero3d_batch *b, int slot, float *r, float *ru, float *rv, float *rw, float *E);
/* Density residual of each listed case (CFD3.readback, src/air_cfd.js:553-554)
 * since its previous call; NaN on a case's first call. Waits. */
int aero3d_batch_residuals(aero3d_batch *b, const int *slots, int n, double *res);
/* cfdForces of one case on the device (src/air_cfd.js:443-452). Waits. */
int aero3d_batch_forces(aero3d_batch *b, int slot, double alphaDeg, double M, double S, double xcg, double zcg, double cref,
                        double out[3]);
int aero3d_batch_footprint(const aero3d_batch *b, int64_t *used, int64_t *capacity);
/* n iterations of every resident case bracketed by HIP events on the batch stream. */
int aero3d_batch_time_steps(aero3d_batch *b, int n, float *ms);

/* ---------------------------------------------------------------------------
 * Served form (no reference counterpart: the reference runs one solve at a
 * time from its UI loop, src/air_cfd.js:539-578). Requests queue FIFO and are
 * admitted into a batch as slots and arena memory free up (admission by HBM
 * footprint, continuous batching). A tick enqueues, without waiting on the
 * stream: the iterations to the next readback of any case, then the readback
 * of every case due (density residual since its previous readback and force
 * coefficients, src/air_cfd.js:552-558) and its convergence rule, decided on
 * the device; a finished case is frozen on the device at once. A case's first
 * readback only takes the density snapshot (residual NaN), as the reference
 * controller's first readback does. The host lags one tick: each tick processes the
 * results of the tick before the last, retires the frozen cases with their
 * coefficients and frees their memory. Readbacks come every check_every
 * iterations from a case's admission (and at max_iters under the residual-
 * drop rule), so a case's iterations, residuals and coefficients do not
 * depend on the other cases: they equal a plain solve read back at the same
 * iterations (aero3d_readback) bit for bit.
 * ------------------------------------------------------------------------- */
#define AERO3D_RULE_RES_DROP 0   /* done at res / res0 <= res_drop, or at max_iters iterations               */
#define AERO3D_RULE_REFERENCE 1  /* the reference controller (src/air_cfd.js:573-576): more than 6 readbacks,
                                    res / res0 < 2e-3 and CL of the last 4 readbacks within 2e-3; or more than
                                    max_iters iterations (the reference uses 30000)                         */
typedef struct aero3d_server aero3d_server;
typedef struct {
  int geometry;                  /* gid registered on aero3d_server_batch(s)        */
  double F[5], fix[15];          /* freestream and fixed states                      */
  double cfl;
  int order;
  double alpha, M, S, xcg, zcg, cref;   /* force-coefficient reference (cfdForces)  */
  int rule;                      /* AERO3D_RULE_*                                    */
  double res_drop;               /* AERO3D_RULE_RES_DROP: res / res0 <= res_drop     */
  int max_iters;                 /* iteration cap                                    */
  double arrival;                /* arrival time on the server clock (s)             */
  int64_t tag;                   /* caller's id, returned in the result              */
  float *state_out;              /* optional: 5 planes of N floats that receive the
                                    final state at retirement (costs a wait then)    */
} aero3d_request;
typedef struct {
  int64_t tag;
  int iters, converged, diverged, failed, nreads;  /* failed: never ran (not admissible) */
  double res0, res, CL, CD, Cm;
  double arrival, t_admit, t_done;  /* server clock (s); latency = t_done - arrival */
} aero3d_result;
int aero3d_server_create(int64_t arena_bytes, int max_slots, int check_every, aero3d_server **out);
int aero3d_server_destroy(aero3d_server *s);
/* The server's batch: register geometries on it with aero3d_batch_add_geometry. */
aero3d_batch *aero3d_server_batch(aero3d_server *s);
/* Seconds since the server was created (steady clock). */
int aero3d_server_clock(const aero3d_server *s, double *now);
/* Queue a request (FIFO). AERO3D_E_ARG for a bad request (geometry, cfl,
 * order, rule, res_drop, max_iters, coefficient reference); AERO3D_E_STATE
 * when its case cannot fit the arena even with no case resident (message on
 * aero3d_server_batch(s)). */
int aero3d_server_submit(aero3d_server *s, const aero3d_request *r);
/* Process the results of finished ticks (waiting only when two ticks are
 * queued), admit, enqueue the next tick; returns resident (not yet retired)
 * and queued counts. A queued request that cannot be admitted with no case
 * resident, or whose geometry was removed, is returned as a result with
 * failed = 1. */
int aero3d_server_tick(aero3d_server *s, int *n_active, int *n_pending);
/* Move up to max finished results to out; *n = number moved. */
int aero3d_server_poll(aero3d_server *s, aero3d_result *out, int max, int *n);
/* Wall time with resident work, device time of the ticks (HIP events around
 * each tick's work), and cell-updates of retired cases. */
int aero3d_server_stats(const aero3d_server *s, double *busy_s, double *gpu_s, int64_t *cell_updates);

#ifdef __cplusplus
}
#endif
#endif
