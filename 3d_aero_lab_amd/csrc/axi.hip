// ============================================================================
// axi.hip — axisymmetric engine solver on gfx950 (C ABI: include/axi.h).
//
// Scheme of the reference CPUSolver (src/cfd_cpu.js:7-210), FP32:
//   face kernel  one thread per face; x faces ((W+1) x H) then r faces
//                (W x (H+1)); primitives of the 2-4 cells a face reads,
//                MUSCL-minmod on full cells (phi > 0.999), HLLC with per-side
//                gamma, Riemann-invariant farfield at x = lo, x = hi, r = top;
//                r faces rotate (normal v, tangential u) and add the
//                mixing-length shear, heat and species fluxes; each face is
//                evaluated once and stored (5 planes) with its eddy viscosity.
//   cell kernel  one thread per cell; flux divergence with r-weighted faces,
//                wall / geometric pressure source, blade-row work and the
//                relaxed heat source, local dt (stage 1), RK update, floors.
//   fused kernel one block of 512 threads per 63 x 7 tile: the tile's
//                primitives (with a 2-cell ring) and face fluxes in LDS, then
//                the cell updates;
//                the same arithmetic as face + cell, one launch per stage.
// Per SSP-RK2 iteration: fused(Q -> Q1, dt), fused(Q1, Q -> Q); with the
// global / dtRatio dt caps face(Q), cell(Q, dt only), dt minimum, cap,
// cell(Q -> Q1, capped dt), face(Q1), cell(Q1, Q -> Q).
// Device layout: Q and Q1 as 5 planes of N floats, face fluxes Fx 5 x
// (W+1)H, Fr 5 x W(H+1), face eddy viscosity W(H+1), dt / src / mfx N floats.
// ============================================================================
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "axi.h"

#include "axi_kernels.hpp"
#include "axi_post.hpp"

using namespace axik;

namespace {

std::string g_axi_err;

}  // namespace

struct axi_solver {
  int W = 0, H = 0, N = 0, it = 0;
  hipStream_t stream = nullptr;
  AxiConsts c;
  float *rc = nullptr, *dx = nullptr, *dr = nullptr, *rf = nullptr;
  uint8_t *solid = nullptr, *comp = nullptr;
  float *zf = nullptr;            // 9 zone planes: phi ae an wgt lmix tgt rate yrate mfd
  float *Q = nullptr, *Q1 = nullptr, *Fx = nullptr, *Fr = nullptr, *nutF = nullptr, *dt = nullptr, *src = nullptr, *mfx = nullptr;
  float *partial = nullptr, *cap = nullptr;
  AxiConsts *dc = nullptr;        // device copy of the constants (read by every kernel)
  bool dc_dirty = true;
  hipGraph_t graph = nullptr;     // captured chunk of kGraphIters iterations
  hipGraphExec_t exec = nullptr;
  bool graph_capped = false;
  int fused = 0;                  // uncapped steps: face + cell kernels (0) or a fused tile kernel form (1 .. 4)
  int graph_fused = -1;
  int global = 0;
  double dtRatio = 0;
  // readback monitors (axi_monitor): FP64 grid widths, the density of the last
  // monitored state, section columns and row masks, partial and final sums
  double *rcD = nullptr, *drD = nullptr, *dxD = nullptr;
  float *rsnap = nullptr;
  bool snap_valid = false;
  int nsec = 0;
  int *sec_col = nullptr;
  uint8_t *sec_rows = nullptr;
  double *mon_part = nullptr, *mon_out = nullptr;
  MonIn *mon_in = nullptr;
  // post-processing: FP64 x centres, the display zone codes, a scratch buffer grown on demand
  double *xcD = nullptr;
  int8_t *dstream = nullptr;
  uint8_t *dmat = nullptr;
  char *post = nullptr;
  size_t post_cap = 0;
  std::string err;
};

namespace {

int afail(axi_solver *s, int code, const std::string &m) { if (s) s->err = m; else g_axi_err = m; return code; }
#define ACHK(s, expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return afail((s), -1000 - (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

AxiView view(const axi_solver *s) {
  AxiView v;
  v.W = s->W; v.H = s->H; v.N = s->N;
  v.rc = s->rc; v.dx = s->dx; v.dr = s->dr; v.rf = s->rf;
  v.solid = s->solid; v.comp = s->comp;
  const size_t N = s->N;
  v.phi = s->zf; v.ae = s->zf + N; v.an = s->zf + 2 * N; v.wgt = s->zf + 3 * N; v.lmix = s->zf + 4 * N;
  v.tgt = s->zf + 5 * N; v.rate = s->zf + 6 * N; v.yrate = s->zf + 7 * N; v.mfd = s->zf + 8 * N;
  return v;
}

int upload_zones(axi_solver *s, const axi_zones *z) {
  if (!z || !z->solid || !z->comp || !z->phi || !z->ae || !z->an || !z->wgt || !z->lmix || !z->tgt || !z->rate || !z->yrate || !z->mfd)
    return afail(s, -1, "zones: null array");
  const size_t N = s->N;
  // component ids index the constant table: checked before anything reaches the device
  for (size_t k = 0; k < N; k++)
    if (z->comp[k] >= 8) return afail(s, -1, "zones: component id >= 8");
  ACHK(s, hipMemcpy(s->solid, z->solid, N, hipMemcpyHostToDevice));
  ACHK(s, hipMemcpy(s->comp, z->comp, N, hipMemcpyHostToDevice));
  const float *planes[9] = {z->phi, z->ae, z->an, z->wgt, z->lmix, z->tgt, z->rate, z->yrate, z->mfd};
  for (int p = 0; p < 9; p++) ACHK(s, hipMemcpy(s->zf + p * N, planes[p], 4 * N, hipMemcpyHostToDevice));
  return 0;
}

constexpr int kGraphIters = 25;

// Both stages of one iteration with the fused kernel of tile height TH and NW waves per block.
template <int TH, int NW>
void fused_iteration(axi_solver *s, const AxiView &v, const AxiConsts *c, int ntx) {
  const int nb = ntx * ((s->H + TH - 1) / TH);
  fused_kernel<0, TH, NW><<<nb, 64 * NW, 0, s->stream>>>(v, c, s->Q, s->Q, s->Q1, s->dt, s->src, s->mfx, ntx);
  fused_kernel<1, TH, NW><<<nb, 64 * NW, 0, s->stream>>>(v, c, s->Q1, s->Q, s->Q, s->dt, s->src, s->mfx, ntx);
}
// fused form f: tiles of 63 x (3, 7, 9, 5) cells, blocks of (4, 8, 16, 4) waves
void launch_fused(axi_solver *s, const AxiView &v, const AxiConsts *c, int ntx) {
  switch (s->fused) {
    case 1: fused_iteration<3, 4>(s, v, c, ntx); break;
    case 2: fused_iteration<7, 8>(s, v, c, ntx); break;
    case 3: fused_iteration<9, 16>(s, v, c, ntx); break;
    default: fused_iteration<5, 4>(s, v, c, ntx); break;
  }
}

void enqueue_iters(axi_solver *s, int n) {
  const AxiView v = view(s);
  const int NF = (s->W + 1) * s->H + s->W * (s->H + 1);
  const dim3 gf((NF + kB - 1) / kB), gc((s->N + kB - 1) / kB);
  const bool capped = s->global || s->dtRatio != 0;
  const int nbm = std::min(1024, (s->N + kB - 1) / kB);
  const AxiConsts *c = s->dc;
  const int ntx = (s->W + kTW - 1) / kTW;
  for (int k = 0; k < n; k++) {
    if (!capped && s->fused) {     // one launch per stage, face fluxes in LDS
      launch_fused(s, v, c, ntx);
      continue;
    }
    face_kernel<<<gf, kB, 0, s->stream>>>(v, c, s->Q, s->Fx, s->Fr, s->nutF);
    if (capped) {
      cell_kernel<<<gc, kB, 0, s->stream>>>(v, c, s->Q, s->Q, s->Q1, s->Fx, s->Fr, s->nutF, s->dt, s->src, s->mfx, 2, s->cap);
      dtmin_kernel<<<nbm, kB, 0, s->stream>>>(s->dt, s->solid, s->N, s->partial);
      dtcap_kernel<<<1, 64, 0, s->stream>>>(s->partial, nbm, c, s->cap);
      cell_kernel<<<gc, kB, 0, s->stream>>>(v, c, s->Q, s->Q, s->Q1, s->Fx, s->Fr, s->nutF, s->dt, s->src, s->mfx, 3, s->cap);
    } else {
      cell_kernel<<<gc, kB, 0, s->stream>>>(v, c, s->Q, s->Q, s->Q1, s->Fx, s->Fr, s->nutF, s->dt, s->src, s->mfx, 0, s->cap);
    }
    face_kernel<<<gf, kB, 0, s->stream>>>(v, c, s->Q1, s->Fx, s->Fr, s->nutF);
    cell_kernel<<<gc, kB, 0, s->stream>>>(v, c, s->Q1, s->Q, s->Q, s->Fx, s->Fr, s->nutF, s->dt, s->src, s->mfx, 1, s->cap);
  }
}

// n iterations: the constants go to the device first (stream-ordered), then
// whole chunks replay a captured graph of kGraphIters iterations (the kernel
// arguments never change: Q and Q1 keep their roles, constants are read from
// device memory) and the remainder is launched directly.
int steps(axi_solver *s, int n) {
  if (s->dc_dirty) {
    ACHK(s, hipMemcpyAsync(s->dc, &s->c, sizeof(AxiConsts), hipMemcpyHostToDevice, s->stream));
    ACHK(s, hipStreamSynchronize(s->stream));   // the host copy of the constants may change after return
    s->dc_dirty = false;
  }
  const bool capped = s->global || s->dtRatio != 0;
  if (s->exec && (s->graph_capped != capped || s->graph_fused != s->fused)) {
    (void)hipGraphExecDestroy(s->exec); (void)hipGraphDestroy(s->graph);
    s->exec = nullptr; s->graph = nullptr;
  }
  int k = 0;
  if (n >= kGraphIters) {
    if (!s->exec) {
      ACHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      enqueue_iters(s, kGraphIters);
      ACHK(s, hipStreamEndCapture(s->stream, &s->graph));
      ACHK(s, hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
      s->graph_capped = capped;
      s->graph_fused = s->fused;
    }
    for (; k + kGraphIters <= n; k += kGraphIters) ACHK(s, hipGraphLaunch(s->exec, s->stream));
  }
  if (k < n) enqueue_iters(s, n - k);
  ACHK(s, hipGetLastError());
  s->it += n;
  return 0;
}

}  // namespace

extern "C" {

int axi_create(const axi_grid *g, const axi_zones *z, const double comps[32], const double fs[3], const double *Q0, axi_solver **out) {
  if (!out) return afail(nullptr, -1, "axi_create: null output");
  *out = nullptr;
  if (!g || !g->xf || !g->rf || !z || !comps || !fs || !Q0 || g->W < 2 || g->H < 2 || (int64_t)g->W * g->H > (int64_t)1 << 28)
    return afail(nullptr, -1, "axi_create: bad argument (W, H >= 2, W*H <= 2^28)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return afail(nullptr, -2, "axi_create: no HIP device");
  axi_solver *s = new axi_solver();
  s->W = g->W; s->H = g->H; s->N = g->W * g->H;
  const size_t N = s->N, NXF = (size_t)(s->W + 1) * s->H, NRF = (size_t)s->W * (s->H + 1);
  auto bail = [&](int rc) { g_axi_err = s->err; axi_destroy(s); return rc; };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return bail(afail(s, -2, "stream"));
  std::memset(&s->c, 0, sizeof s->c);
  s->c.cfl = 0.6f; s->c.order = 2; s->c.visc = 1; s->c.hf = 0.99f * 43e6f;
  for (int q = 0; q < 32; q++) s->c.comps[q] = (float)comps[q];
  s->c.fs_r = (float)fs[0]; s->c.fs_u = (float)fs[1]; s->c.fs_p = (float)fs[2];
  std::vector<float> rc(s->H), dx(s->W), dr(s->H), rf(s->H + 1);
  for (int i = 0; i < s->W; i++) dx[i] = (float)(g->xf[i + 1] - g->xf[i]);
  for (int j = 0; j < s->H; j++) { rc[j] = (float)(0.5 * (g->rf[j] + g->rf[j + 1])); dr[j] = (float)(g->rf[j + 1] - g->rf[j]); }
  for (int j = 0; j <= s->H; j++) rf[j] = (float)g->rf[j];
  auto al = [&](auto **p, size_t n) { return hipMalloc((void **)p, sizeof(**p) * (n ? n : 1)); };
  if (al(&s->rc, s->H) || al(&s->dx, s->W) || al(&s->dr, s->H) || al(&s->rf, s->H + 1) || al(&s->solid, N) || al(&s->comp, N) ||
      al(&s->zf, 9 * N) || al(&s->Q, 5 * N) || al(&s->Q1, 5 * N) || al(&s->Fx, 5 * NXF) || al(&s->Fr, 5 * NRF) || al(&s->nutF, NRF) ||
      al(&s->dt, N) || al(&s->src, N) || al(&s->mfx, N) || al(&s->partial, 1024) || al(&s->cap, 1) || al(&s->dc, 1) ||
      al(&s->rcD, s->H) || al(&s->drD, s->H) || al(&s->dxD, s->W) || al(&s->rsnap, N) || al(&s->sec_col, kMonMaxSec) ||
      al(&s->sec_rows, (size_t)kMonMaxSec * s->H) || al(&s->mon_part, (size_t)kMonBlocks * kMonK) ||
      al(&s->mon_out, kMonK + kSecK * kMonMaxSec) || al(&s->mon_in, 1) || al(&s->xcD, s->W))
    return bail(afail(s, -1002, "axi_create: hipMalloc failed"));
  if (hipMemcpy(s->rc, rc.data(), 4 * rc.size(), hipMemcpyHostToDevice) || hipMemcpy(s->dx, dx.data(), 4 * dx.size(), hipMemcpyHostToDevice) ||
      hipMemcpy(s->dr, dr.data(), 4 * dr.size(), hipMemcpyHostToDevice) || hipMemcpy(s->rf, rf.data(), 4 * rf.size(), hipMemcpyHostToDevice) ||
      hipMemset(s->src, 0, 4 * N) || hipMemset(s->mfx, 0, 4 * N) || hipMemset(s->dt, 0, 4 * N))
    return bail(afail(s, -1001, "axi_create: upload failed"));
  {  // the host grid's double centres and widths (buildCFDGrid, src/cfd_grid.js:36-41)
    std::vector<double> rcD(s->H), drD(s->H), dxD(s->W), xcD(s->W);
    for (int i = 0; i < s->W; i++) { dxD[i] = g->xf[i + 1] - g->xf[i]; xcD[i] = 0.5 * (g->xf[i] + g->xf[i + 1]); }
    for (int j = 0; j < s->H; j++) { rcD[j] = 0.5 * (g->rf[j] + g->rf[j + 1]); drD[j] = g->rf[j + 1] - g->rf[j]; }
    if (hipMemcpy(s->rcD, rcD.data(), 8 * rcD.size(), hipMemcpyHostToDevice) || hipMemcpy(s->drD, drD.data(), 8 * drD.size(), hipMemcpyHostToDevice) ||
        hipMemcpy(s->dxD, dxD.data(), 8 * dxD.size(), hipMemcpyHostToDevice) || hipMemcpy(s->xcD, xcD.data(), 8 * xcD.size(), hipMemcpyHostToDevice))
      return bail(afail(s, -1001, "axi_create: upload failed"));
  }
  {  // default kernel form: tiles of 63 x 9 cells in blocks of 16 waves when they make between half and all
     // of one block per CU (one round, each CU's block spread over its four SIMDs), else face + cell kernels
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0) {
      const long long b3 = (long long)((s->W + kTW - 1) / kTW) * ((s->H + 8) / 9);
      s->fused = (2 * b3 >= ncu && b3 <= ncu) ? 3 : 0;
    }
  }
  int rc_;
  if ((rc_ = upload_zones(s, z))) return bail(rc_);
  if ((rc_ = axi_set_state(s, Q0))) return bail(rc_);
  *out = s;
  return 0;
}

int axi_destroy(axi_solver *s) {
  if (!s) return 0;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  void *p[] = {s->rc, s->dx, s->dr, s->rf, s->solid, s->comp, s->zf, s->Q, s->Q1, s->Fx, s->Fr, s->nutF, s->dt, s->src, s->mfx, s->partial, s->cap, s->dc,
               s->rcD, s->drD, s->dxD, s->rsnap, s->sec_col, s->sec_rows, s->mon_part, s->mon_out, s->mon_in, s->xcD, s->dstream, s->dmat,
               s->post};
  for (void *x : p) if (x) (void)hipFree(x);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

const char *axi_last_error(const axi_solver *s) { return s ? s->err.c_str() : g_axi_err.c_str(); }

int axi_set_opt(axi_solver *s, double cfl, int order, int visc, double hf, int global, double dtRatio) {
  if (!s) return -1;
  if (!(cfl > 0 && cfl < 10) || (order != 1 && order != 2) || dtRatio < 0) return afail(s, -1, "set_opt: cfl in (0,10), order 1|2, dtRatio >= 0");
  s->c.cfl = (float)cfl; s->c.order = order; s->c.visc = visc ? 1 : 0; s->c.hf = hf > 0 ? (float)hf : 0.99f * 43e6f;
  s->global = global ? 1 : 0; s->dtRatio = dtRatio;
  s->c.capf = s->global ? 1.f : (float)dtRatio;
  s->dc_dirty = true;
  return 0;
}
int axi_set_fused(axi_solver *s, int on) {
  if (!s || on < 0 || on > 4) return afail(s, -1, "set_fused: 0 .. 4");
  s->fused = on;                 // the next step re-captures its graph
  return 0;
}
int axi_set_comps(axi_solver *s, const double comps[32]) {
  if (!s || !comps) return -1;
  for (int q = 0; q < 32; q++) s->c.comps[q] = (float)comps[q];
  s->dc_dirty = true;
  return 0;
}
int axi_set_zones(axi_solver *s, const axi_zones *z) {
  if (!s) return -1;
  ACHK(s, hipStreamSynchronize(s->stream));
  return upload_zones(s, z);
}
int axi_set_fs(axi_solver *s, const double fs[3]) {
  if (!s || !fs) return -1;
  s->c.fs_r = (float)fs[0]; s->c.fs_u = (float)fs[1]; s->c.fs_p = (float)fs[2];
  s->dc_dirty = true;
  return 0;
}
int axi_set_state(axi_solver *s, const double *Q) {
  if (!s || !Q) return -1;
  std::vector<float> f(5 * (size_t)s->N);
  for (size_t k = 0; k < f.size(); k++) f[k] = (float)Q[k];
  ACHK(s, hipStreamSynchronize(s->stream));
  ACHK(s, hipMemcpy(s->Q, f.data(), 4 * f.size(), hipMemcpyHostToDevice));
  s->snap_valid = false;          // the next monitor has no previous readback
  return 0;
}
int axi_step(axi_solver *s, int n) {
  if (!s || n < 0) return afail(s, -1, "step: n >= 0");
  return steps(s, n);
}
int axi_iter(const axi_solver *s, int *it) {
  if (!s || !it) return -1;
  *it = s->it;
  return 0;
}
int axi_read(axi_solver *s, float *Q, float *src, float *mfx, float *dt) {
  if (!s) return -1;
  const size_t b = 4 * (size_t)s->N;
  if (Q) ACHK(s, hipMemcpyAsync(Q, s->Q, 5 * b, hipMemcpyDeviceToHost, s->stream));
  if (src) ACHK(s, hipMemcpyAsync(src, s->src, b, hipMemcpyDeviceToHost, s->stream));
  if (mfx) ACHK(s, hipMemcpyAsync(mfx, s->mfx, b, hipMemcpyDeviceToHost, s->stream));
  if (dt) ACHK(s, hipMemcpyAsync(dt, s->dt, b, hipMemcpyDeviceToHost, s->stream));
  ACHK(s, hipStreamSynchronize(s->stream));
  return 0;
}
int axi_monitor_sections(axi_solver *s, int nsec, const int *col, const uint8_t *rows) {
  if (!s || nsec < 0 || nsec > kMonMaxSec || (nsec > 0 && (!col || !rows))) return afail(s, -1, "monitor_sections: bad argument");
  for (int q = 0; q < nsec; q++)
    if (col[q] < 0 || col[q] >= s->W) return afail(s, -1, "monitor_sections: column out of range");
  ACHK(s, hipStreamSynchronize(s->stream));
  if (nsec > 0) {
    ACHK(s, hipMemcpy(s->sec_col, col, sizeof(int) * nsec, hipMemcpyHostToDevice));
    ACHK(s, hipMemcpy(s->sec_rows, rows, (size_t)nsec * s->H, hipMemcpyHostToDevice));
  }
  s->nsec = nsec;
  return 0;
}

int axi_monitor(axi_solver *s, const double comps[32], double P0, axi_monitor_out *out) {
  if (!s || !comps || !out) return afail(s, -1, "monitor: null argument");
  MonIn in;
  std::memcpy(in.comps, comps, sizeof in.comps);
  in.P0 = P0;
  in.prev_valid = s->snap_valid ? 1 : 0;
  ACHK(s, hipMemcpyAsync(s->mon_in, &in, sizeof in, hipMemcpyHostToDevice, s->stream));
  const AxiView v = view(s);
  const int nb = std::max(1, std::min(kMonBlocks, (s->N + kB - 1) / kB));
  monitor_kernel<<<nb, kB, 0, s->stream>>>(v, s->rcD, s->drD, s->dxD, s->Q, s->src, s->mfx, s->rsnap, s->mon_in, s->mon_part);
  monitor_finish<<<1, kMonFinishThreads, 0, s->stream>>>(s->mon_part, nb, s->mon_out);
  if (s->nsec > 0)
    section_kernel<<<s->nsec, kSecBlock, 0, s->stream>>>(v, s->rcD, s->drD, s->Q, s->mfx, s->nsec, s->sec_col, s->sec_rows, s->mon_out + kMonK);
  ACHK(s, hipGetLastError());
  double o[kMonK + kSecK * kMonMaxSec];
  ACHK(s, hipMemcpyAsync(o, s->mon_out, sizeof(double) * (kMonK + kSecK * s->nsec), hipMemcpyDeviceToHost, s->stream));
  ACHK(s, hipMemcpyAsync(s->rsnap, s->Q, 4 * (size_t)s->N, hipMemcpyDeviceToDevice, s->stream));   // density of this readback
  ACHK(s, hipStreamSynchronize(s->stream));
  s->snap_valid = true;
  std::memset(out, 0, sizeof *out);
  out->Fb = o[0]; out->Fw = o[1]; out->F = o[0] + o[1]; out->mfuel = o[2];
  for (int q = 0; q < 8; q++) out->pw[q] = o[3 + q];
  out->res_sum2 = o[11]; out->res_n = o[12]; out->res_valid = in.prev_valid;
  out->bad = (int)o[13];
  out->nsec = s->nsec;
  for (int q = 0; q < s->nsec; q++) {
    const double *sq = o + kMonK + kSecK * q;
    out->sec_mdot[q] = sq[0]; out->sec_w[q] = sq[1]; out->sec_wTt[q] = sq[4];
  }
  return 0;
}

// ---------------------------------------------------------------------------
// post-processing on the device (section averages, display fields, transfer)
// ---------------------------------------------------------------------------
namespace {
constexpr int kMaxSections = 256;

// the post scratch buffer, at least `bytes` (grown after the stream drained)
int post_buf(axi_solver *s, size_t bytes) {
  if (bytes <= s->post_cap) return 0;
  ACHK(s, hipStreamSynchronize(s->stream));
  if (s->post) { ACHK(s, hipFree(s->post)); s->post = nullptr; s->post_cap = 0; }
  ACHK(s, hipMalloc((void **)&s->post, bytes));
  s->post_cap = bytes;
  return 0;
}
size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

int axi_sections(axi_solver *s, int nsec, const int *col, const uint8_t *rows, double *sums) {
  if (!s || nsec < 0 || nsec > kMaxSections || (nsec > 0 && (!col || !rows || !sums)))
    return afail(s, -1, "sections: bad argument (0 <= nsec <= 256)");
  for (int q = 0; q < nsec; q++)
    if (col[q] < 0 || col[q] >= s->W) return afail(s, -1, "sections: column out of range");
  if (nsec == 0) return 0;
  const size_t bc = up256(sizeof(int) * nsec), br = up256((size_t)nsec * s->H), bo = sizeof(double) * kSecK * nsec;
  if (int rc = post_buf(s, bc + br + bo)) return rc;
  int *dcol = (int *)s->post;
  uint8_t *drows = (uint8_t *)(s->post + bc);
  double *dout = (double *)(s->post + bc + br);
  ACHK(s, hipMemcpyAsync(dcol, col, sizeof(int) * nsec, hipMemcpyHostToDevice, s->stream));
  ACHK(s, hipMemcpyAsync(drows, rows, (size_t)nsec * s->H, hipMemcpyHostToDevice, s->stream));
  section_kernel<<<nsec, kSecBlock, 0, s->stream>>>(view(s), s->rcD, s->drD, s->Q, s->mfx, nsec, dcol, drows, dout);
  ACHK(s, hipGetLastError());
  ACHK(s, hipMemcpyAsync(sums, dout, bo, hipMemcpyDeviceToHost, s->stream));
  ACHK(s, hipStreamSynchronize(s->stream));
  return 0;
}

int axi_set_display_zones(axi_solver *s, const int8_t *stream, const uint8_t *mat) {
  if (!s || !stream || !mat) return afail(s, -1, "set_display_zones: null argument");
  const size_t N = s->N;
  if (!s->dstream) {
    ACHK(s, hipMalloc((void **)&s->dstream, N));
    ACHK(s, hipMalloc((void **)&s->dmat, N));
  }
  ACHK(s, hipMemcpyAsync(s->dstream, stream, N, hipMemcpyHostToDevice, s->stream));
  ACHK(s, hipMemcpyAsync(s->dmat, mat, N, hipMemcpyHostToDevice, s->stream));
  ACHK(s, hipStreamSynchronize(s->stream));
  return 0;
}

int axi_derived(axi_solver *s, int disp, double T0, double P0, float *fields, uint8_t *reg, uint8_t *mat) {
  if (!s || (disp != 0 && disp != 1)) return afail(s, -1, "derived: bad argument (disp 0 or 1)");
  if (!s->dstream) return afail(s, -3, "derived: axi_set_display_zones first");
  const size_t N = s->N, bf = up256(sizeof(float) * kDerivedFields * N), bb = up256(N);
  if (int rc = post_buf(s, bf + 2 * bb)) return rc;
  float *df = (float *)s->post;
  uint8_t *dreg = (uint8_t *)(s->post + bf), *dmat = (uint8_t *)(s->post + bf + bb);
  derived_kernel<<<(s->N + kB - 1) / kB, kB, 0, s->stream>>>(view(s), s->xcD, s->rcD, s->Q, s->dstream, s->dmat, disp, T0, P0, df, dreg,
                                                              dmat);
  ACHK(s, hipGetLastError());
  if (fields) ACHK(s, hipMemcpyAsync(fields, df, sizeof(float) * kDerivedFields * N, hipMemcpyDeviceToHost, s->stream));
  if (reg) ACHK(s, hipMemcpyAsync(reg, dreg, N, hipMemcpyDeviceToHost, s->stream));
  if (mat) ACHK(s, hipMemcpyAsync(mat, dmat, N, hipMemcpyDeviceToHost, s->stream));
  ACHK(s, hipStreamSynchronize(s->stream));
  return 0;
}

int axi_transfer(axi_solver *dst, const axi_solver *src, double T0, double P0) {
  if (!dst || !src || dst == src) return afail(dst, -1, "transfer: two different solvers");
  hipPointerAttribute_t a{}, b{};
  if (hipPointerGetAttributes(&a, src->Q) != hipSuccess || hipPointerGetAttributes(&b, dst->Q) != hipSuccess || a.device != b.device)
    return afail(dst, -1, "transfer: the solvers are on different devices");
  ACHK(dst, hipStreamSynchronize(src->stream));   // the source's steps have run
  InterpSrc A{src->W, src->H, src->xcD, src->rcD, src->solid, src->Q};
  interp_kernel<<<(dst->N + kB - 1) / kB, kB, 0, dst->stream>>>(A, dst->W, dst->H, dst->xcD, dst->rcD, T0, P0, dst->Q);
  ACHK(dst, hipGetLastError());
  ACHK(dst, hipStreamSynchronize(dst->stream));
  dst->snap_valid = false;        // as after setState: the next monitor has no previous readback
  return 0;
}

int axi_sync(axi_solver *s) {
  if (!s) return -1;
  ACHK(s, hipStreamSynchronize(s->stream));
  return 0;
}
int axi_autotune(axi_solver *s, int iters, int *form, float *ms_best) {
  if (!s || iters < 1) return afail(s, -1, "autotune: iters >= 1");
  // time on a copy: save the state and the diagnostics, restore after
  const size_t b = 4 * (size_t)s->N;
  float *save = nullptr;
  ACHK(s, hipMalloc((void **)&save, 8 * b));
  struct Free { float *p; ~Free() { if (p) (void)hipFree(p); } } fr{save};
  ACHK(s, hipMemcpyAsync(save, s->Q, 5 * b, hipMemcpyDeviceToDevice, s->stream));
  ACHK(s, hipMemcpyAsync(save + 5 * (size_t)s->N, s->dt, b, hipMemcpyDeviceToDevice, s->stream));
  ACHK(s, hipMemcpyAsync(save + 6 * (size_t)s->N, s->src, b, hipMemcpyDeviceToDevice, s->stream));
  ACHK(s, hipMemcpyAsync(save + 7 * (size_t)s->N, s->mfx, b, hipMemcpyDeviceToDevice, s->stream));
  const int it0 = s->it, f0 = s->fused;
  int best = f0;
  float tbest = 1e30f;
  for (int f = 0; f <= 4; f++) {
    s->fused = f;
    float ms;
    int rc = axi_time_steps(s, kGraphIters, &ms);   // warm (and capture)
    if (!rc) rc = axi_time_steps(s, iters, &ms);
    if (rc) { s->fused = f0; return rc; }
    if (ms < tbest) { tbest = ms; best = f; }
  }
  s->fused = best;
  s->it = it0;
  ACHK(s, hipMemcpyAsync(s->Q, save, 5 * b, hipMemcpyDeviceToDevice, s->stream));
  ACHK(s, hipMemcpyAsync(s->dt, save + 5 * (size_t)s->N, b, hipMemcpyDeviceToDevice, s->stream));
  ACHK(s, hipMemcpyAsync(s->src, save + 6 * (size_t)s->N, b, hipMemcpyDeviceToDevice, s->stream));
  ACHK(s, hipMemcpyAsync(s->mfx, save + 7 * (size_t)s->N, b, hipMemcpyDeviceToDevice, s->stream));
  ACHK(s, hipStreamSynchronize(s->stream));
  if (form) *form = best;
  if (ms_best) *ms_best = tbest / iters;
  return 0;
}

int axi_time_steps(axi_solver *s, int n, float *ms) {
  if (!s || !ms || n < 0) return -1;
  struct Ev { hipEvent_t e = nullptr; ~Ev() { if (e) (void)hipEventDestroy(e); } } e0, e1;   // released on every path
  ACHK(s, hipEventCreate(&e0.e));
  ACHK(s, hipEventCreate(&e1.e));
  ACHK(s, hipEventRecord(e0.e, s->stream));
  int rc = steps(s, n);
  if (rc) return rc;
  ACHK(s, hipEventRecord(e1.e, s->stream));
  ACHK(s, hipEventSynchronize(e1.e));
  ACHK(s, hipEventElapsedTime(ms, e0.e, e1.e));
  return 0;
}

}  // extern "C"
