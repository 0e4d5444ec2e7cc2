// ============================================================================
// aero3d.hip — libaero3d.so: HIP kernels and C ABI (include/aero3d.h) of the
// plain 3D Euler cut-cell solver on one MI355X (gfx950).
//
// Device layout of one solve (N = NX*NY*NZ cells, cidx order, x fastest):
//   state   3 buffers x 5 planes x ns floats (ns = N rounded up to 64):
//           current Q, stage-1 result S1, stage-2 result B, rotated like the
//           reference's three ping-pong textures (src/air_cfd.js:421-425)
//   masks   t u8[N], phi f32[N], ap f32[4][N]
//   grid    dx, dy, dz f32 widths; cen f32 centres of x, y, z (NX + NY + NZ)
//   dt      f32[N] local time step of the last stage
//   rsnap   f32[N] density snapshot for the residual
//   force   list q i32, cx, cz, x, z f64 and FP64 block partials
// ============================================================================
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "aero3d.h"
#include "euler3d.hpp"
#include "euler3d_tile.hpp"
#include "probe.hpp"
#include "reduce.hpp"

using namespace aero3d;

namespace {

std::string g_create_error;   // message of the last failed aero3d_create

constexpr int kBlock = 256;
constexpr int kProbeMaxPoints = 1 << 26;

}  // namespace

struct aero3d_solver {
  int NX = 0, NY = 0, NZ = 0, N = 0;
  size_t ns = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  float *dx = nullptr, *dy = nullptr, *dz = nullptr;
  float *cen = nullptr;       // FP32 centres: x (NX), then y (NY), then z (NZ)
  std::vector<float> hxc, hdy, hdz;
  uint8_t *t = nullptr;
  float *phi = nullptr, *ap = nullptr;
  float *state = nullptr;     // 3 * 5 * ns
  float *dt = nullptr, *rsnap = nullptr;
  int cur = 0;                // buffer index holding the current state Q
  int tile_ty = 8, tile_zc = 0; // stage kernel: ty 0 = per-cell kernel, 4 / 8: 64 x ty tiles, 32: 32 x 8 tiles; zc > 0 fixed z chunks,
                                //   zc <= 0 balanced launch of (max(1, -zc)) x resident blocks
  int resident4 = 0, resident8 = 0, resident32 = 0; // resident blocks of the balanced kernels on this device
  int it = 0, snap_it = -1;
  double F[5] = {1, 0, 0, 0, 1 / (1.4 * 0.4)};
  double fix[3][5];
  double cfl = 0.7;
  int order = 2;
  // force list
  int nf = 0;
  int32_t *fq = nullptr;
  double *fcx = nullptr, *fcz = nullptr, *fx = nullptr, *fz = nullptr;
  double *partials = nullptr, *sums = nullptr;
  RbSeg *dseg = nullptr;      // device copy of this solve's readback segment
  char *probe = nullptr;      // aero3d_probe points and results, grown on demand
  size_t probe_cap = 0;
  std::string err;
  int64_t bytes = 0;
};

namespace {

int fail(aero3d_solver *s, int code, const std::string &msg) {
  if (s) s->err = msg; else g_create_error = msg;
  return code;
}

#define HIPCHK(s, expr)                                                                   \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail((s), AERO3D_E_HIP_BASE - (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int dalloc(aero3d_solver *s, T **p, size_t n) {
  HIPCHK(s, hipMalloc((void **)p, sizeof(T) * (n ? n : 1)));
  s->bytes += (int64_t)(sizeof(T) * n);
  return 0;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) stage_kernel(CaseView cv, StageConsts kc, const float *__restrict__ qin,
                                                       const float *__restrict__ bse, float *__restrict__ out,
                                                       float *__restrict__ dtout, size_t ns, float wa, float wb) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= cv.N) return;
  stage_cell(cv, kc, qin, bse, out, dtout, ns, wa, wb, q);
}

template <int TX, int TY>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(4))) tile_kernel(CaseView cv, StageConsts kc, const float *__restrict__ qin,
                                                      const float *__restrict__ bse, float *__restrict__ out,
                                                      float *__restrict__ dtout, size_t ns, float wa, float wb, int ntx, int nty,
                                                      int zc) {
  __shared__ TileSmem<TX, TY> sm;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = lb % ntx, by = (lb / ntx) % nty, bz = lb / (ntx * nty);
  const int k0 = bz * zc, k1 = min(cv.NZ, k0 + zc);
  tile_stage<TX, TY>(cv, kc, qin, bse, out, dtout, ns, wa, wb, bx * TX, by * TY, k0, k1, sm);
}

// Balanced launch: every tile gets the same number c of z chunks (c = G / nt
// with G a multiple of the device's resident blocks), chunk l of a tile holds
// planes [l NZ / c, (l+1) NZ / c). Logical block L = l * nt + t is chunk-major
// and the XCD remap hands each XCD a contiguous range of L, i.e. whole chunk
// levels: x-y neighbour tiles (which read each other's halo rows) and the
// chunk below (whose top planes start this chunk's z queue) share an L2.
// kPrio (launches of at most one resident round): wave priority from chunk
// progress, so the two blocks of a CU finish together. kMode: as tile_stage.
template <int TX, int TY, bool kPrio, int kMode>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(4))) tile_kernel_bal(CaseView cv, StageConsts kc, const float *__restrict__ qin,
                                                          const float *__restrict__ bse, float *__restrict__ out,
                                                          float *__restrict__ dtout, size_t ns, float wa, float wb, int ntx, int nty) {
  __shared__ TileSmem<TX, TY> sm;
  const int G = gridDim.x, nt = ntx * nty, c = G / nt;
  const int L = xcd_remap(blockIdx.x, G);
  const int l = L / nt, t = L % nt;
  const int k0 = (int)(((long long)l * cv.NZ) / c), k1 = (int)(((long long)(l + 1) * cv.NZ) / c);
  const int bx = t % ntx, by = t / ntx;
  tile_stage<TX, TY, kPrio, kMode, (TY >= 8)>(cv, kc, qin, bse, out, dtout, ns, wa, wb, bx * TX, by * TY, k0, k1, sm);
}

// Reset: every plane of buffer `dst` := F.
__global__ void fill_kernel(float *__restrict__ dst, size_t ns, int N, float f0, float f1, float f2, float f3, float f4) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= N) return;
  dst[q] = f0; dst[ns + q] = f1; dst[2 * ns + q] = f2; dst[3 * ns + q] = f3; dst[4 * ns + q] = f4;
}

// Fixed cells (flags 3..5) of buffer `dst` := their conserved fixed states. The
// stage kernels take a fixed cell's primitives from the stored state, so every
// state a stage reads holds them (the stages write them too).
struct FixStates { float q[3][5]; };
__global__ void fixcells_kernel(float *__restrict__ dst, size_t ns, int N, const uint8_t *__restrict__ t, FixStates f) {
  const int q = blockIdx.x * kBlock + threadIdx.x;
  if (q >= N) return;
  const int g = t[q] - 3;
  if (g < 0) return;
  for (int c = 0; c < 5; c++) dst[c * ns + q] = f.q[g][c];
}


// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
void prim5_host(const double *a, double *o) {
  const double r = std::max(a[0], 1e-6), u = a[1] / r, v = a[2] / r, w = a[3] / r;
  o[0] = r; o[1] = u; o[2] = v; o[3] = w;
  o[4] = std::max((1.4 - 1) * (a[4] - 0.5 * r * (u * u + v * v + w * w)), 1e-7);
}

StageConsts make_consts(const aero3d_solver *s) {
  StageConsts k;
  double w[5];
  prim5_host(s->F, w);
  for (int v = 0; v < 5; v++) k.Fw[v] = (float)w[v];
  for (int g = 0; g < 3; g++) {
    prim5_host(s->fix[g], w);
    for (int v = 0; v < 5; v++) { k.fixw[g][v] = (float)w[v]; k.fixq[g][v] = (float)s->fix[g][v]; }
  }
  k.cfl = (float)s->cfl;
  k.order = s->order;
  return k;
}

CaseView make_view(const aero3d_solver *s) {
  CaseView cv;
  cv.NX = s->NX; cv.NY = s->NY; cv.NZ = s->NZ; cv.N = s->N;
  cv.dx = s->dx; cv.dy = s->dy; cv.dz = s->dz;
  cv.t = s->t; cv.phi = s->phi; cv.ap = s->ap;
  return cv;
}

inline float *buf(aero3d_solver *s, int b) { return s->state + (size_t)(b % 3) * 5 * s->ns; }

// Write the fixed states into the fixed cells of the current state.
int apply_fix(aero3d_solver *s) {
  FixStates f;
  for (int g = 0; g < 3; g++)
    for (int c = 0; c < 5; c++) f.q[g][c] = (float)s->fix[g][c];
  fixcells_kernel<<<(s->N + kBlock - 1) / kBlock, kBlock, 0, s->stream>>>(buf(s, s->cur), s->ns, s->N, s->t, f);
  HIPCHK(s, hipGetLastError());
  return 0;
}

// The balanced kernels by [one resident round][stage mode - 1]; stage mode 1 =
// stage 1 (qin == bse, which never writes dt), 2 = stage 2, 3 = stage 2 with dt.
using BalKernel = void (*)(CaseView, StageConsts, const float *, const float *, float *, float *, size_t, float, float, int, int);
const BalKernel kBal4[2][3] = {{tile_kernel_bal<64, 4, false, 1>, tile_kernel_bal<64, 4, false, 2>, tile_kernel_bal<64, 4, false, 3>},
                               {tile_kernel_bal<64, 4, true, 1>, tile_kernel_bal<64, 4, true, 2>, tile_kernel_bal<64, 4, true, 3>}};
const BalKernel kBal8[2][3] = {{tile_kernel_bal<64, 8, false, 1>, tile_kernel_bal<64, 8, false, 2>, tile_kernel_bal<64, 8, false, 3>},
                               {tile_kernel_bal<64, 8, true, 1>, tile_kernel_bal<64, 8, true, 2>, tile_kernel_bal<64, 8, true, 3>}};
const BalKernel kBal32[2][3] = {{tile_kernel_bal<32, 8, false, 1>, tile_kernel_bal<32, 8, false, 2>, tile_kernel_bal<32, 8, false, 3>},
                                {tile_kernel_bal<32, 8, true, 1>, tile_kernel_bal<32, 8, true, 2>, tile_kernel_bal<32, 8, true, 3>}};

// Tile shape of a shape code: 4 and 8 are 64 x ty tiles, 32 is 32 x 8 tiles.
int tile_tx(int code) { return code == 32 ? 32 : 64; }
int tile_th(int code) { return code == 32 ? 8 : code; }

void launch_stage(aero3d_solver *s, const CaseView &cv, const StageConsts &kc, const float *qin, const float *bse, float *out,
                  float *dtout, float wa, float wb) {
  if (s->tile_ty == 0) {
    stage_kernel<<<(s->N + kBlock - 1) / kBlock, kBlock, 0, s->stream>>>(cv, kc, qin, bse, out, dtout, s->ns, wa, wb);
    return;
  }
  const int TX = tile_tx(s->tile_ty), TH = tile_th(s->tile_ty);
  const int ntx = (s->NX + TX - 1) / TX, nty = (s->NY + TH - 1) / TH;
  if (s->tile_zc <= 0) {            // balanced launch, G = multiple of the resident block count
    const int mult = s->tile_zc == 0 ? 1 : -s->tile_zc;
    const long long nt = (long long)ntx * nty;
    const int resident = s->tile_ty == 4 ? s->resident4 : (s->tile_ty == 32 ? s->resident32 : s->resident8);
    const long long res = (long long)mult * resident;
    const long long c = std::max(1LL, std::min((long long)s->NZ, res / nt));   // z chunks per tile
    const long long G = nt * c;
    const bool one_round = G <= resident;
    const int mode = qin == bse ? 1 : (dtout ? 3 : 2);
    const BalKernel k = s->tile_ty == 4 ? kBal4[one_round][mode - 1] : (s->tile_ty == 32 ? kBal32[one_round][mode - 1] : kBal8[one_round][mode - 1]);
    k<<<(int)G, dim3(TX, TH), 0, s->stream>>>(cv, kc, qin, bse, out, dtout, s->ns, wa, wb, ntx, nty);
    return;
  }
  const int nzc = (s->NZ + s->tile_zc - 1) / s->tile_zc;
  const int nb = ntx * nty * nzc;
  if (s->tile_ty == 4)
    tile_kernel<64, 4><<<nb, dim3(64, 4), 0, s->stream>>>(cv, kc, qin, bse, out, dtout, s->ns, wa, wb, ntx, nty, s->tile_zc);
  else if (s->tile_ty == 32)
    tile_kernel<32, 8><<<nb, dim3(32, 8), 0, s->stream>>>(cv, kc, qin, bse, out, dtout, s->ns, wa, wb, ntx, nty, s->tile_zc);
  else
    tile_kernel<64, 8><<<nb, dim3(64, 8), 0, s->stream>>>(cv, kc, qin, bse, out, dtout, s->ns, wa, wb, ntx, nty, s->tile_zc);
}

int enqueue_steps(aero3d_solver *s, int n) {
  const CaseView cv = make_view(s);
  const StageConsts kc = make_consts(s);
  for (int k = 0; k < n; k++) {
    float *Q = buf(s, s->cur), *S1 = buf(s, s->cur + 1), *B = buf(s, s->cur + 2);
    launch_stage(s, cv, kc, Q, Q, S1, nullptr, 0.f, 1.f);
    launch_stage(s, cv, kc, S1, Q, B, k == n - 1 ? s->dt : nullptr, 0.5f, 0.5f);
    s->cur = (s->cur + 2) % 3;
    s->it++;
  }
  HIPCHK(s, hipGetLastError());
  return 0;
}


// Readback sums of the current state through the shared segment kernels:
// the residual since the snapshot (then the snapshot is retaken) when `res`,
// the force-list sums and coefficients when `frc`. o[] gets kRbOut values
// (plus the Trefftz sum and squared Mach maximum when `extra`); waits.
int readback_sums(aero3d_solver *s, bool res, bool frc, double alphaDeg, double M, double S, double xcg, double zcg, double cref,
                  double *o, bool extra = false) {
  RbSeg g{};
  g.Q = buf(s, s->cur); g.ns = s->ns; g.t = s->t; g.N = s->N; g.rsnap = s->rsnap;
  g.nit = (s->snap_it >= 0 && s->it > s->snap_it) ? s->it - s->snap_it : 0;
  g.fq = s->fq; g.cx = s->fcx; g.cz = s->fcz; g.x = s->fx; g.z = s->fz; g.nf = frc ? s->nf : 0;
  g.xcg = xcg; g.zcg = zcg;
  const double a = alphaDeg * (M_PI / 180);
  g.ca = std::cos(a); g.sa = std::sin(a); g.kq = 2 / (0.5 * M * M * S); g.cref = cref;
  g.nb_res = res ? reduce_blocks((s->N + 2) / 3) : 0;
  g.nb_f = frc ? reduce_blocks(s->nf) : 0;
  g.blk0 = 0; g.slot = 0; g.it = s->it;
  g.part = s->partials; g.out = s->sums;
  HIPCHK(s, hipMemcpyAsync(s->dseg, &g, sizeof g, hipMemcpyHostToDevice, s->stream));
  if (g.nb_res + g.nb_f > 0) rb_partials_kernel<<<g.nb_res + g.nb_f, kRedBlock, 0, s->stream>>>(s->dseg, 1);
  rb_finish_kernel<<<1, kRedBlock, 0, s->stream>>>(s->dseg);
  HIPCHK(s, hipGetLastError());
  HIPCHK(s, hipMemcpyAsync(o, s->sums, sizeof(double) * (kRbOut + (extra ? 2 : 0)), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  if (res) s->snap_it = s->it;
  if (!res) o[5] = NAN;
  return 0;
}

// Trefftz plane sum at xT into sums[kRbOut] (cfdLocate on the FP32 centres,
// src/air_cfd.js:453, 560-569); enqueued.
int enqueue_trefftz(aero3d_solver *s, double xT, double alphaDeg) {
  const std::vector<float> &c = s->hxc;
  const int n = s->NX;
  int i0; double tx;
  if (xT <= c[0]) { i0 = 0; tx = 0; }
  else if (xT >= c[n - 1]) { i0 = n - 2; tx = 1; }
  else {
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (c[m] <= xT) lo = m; else hi = m; }
    i0 = lo; tx = (xT - c[lo]) / ((double)c[hi] - c[lo]);
  }
  const double a = alphaDeg * (M_PI / 180);
  const int nb = reduce_blocks(s->NY * s->NZ);
  trefftz_kernel<<<nb, kRedBlock, 0, s->stream>>>(buf(s, s->cur), s->ns, s->t, s->dy, s->dz, s->NX, s->NY, s->NZ, i0, tx,
                                                  std::cos(a), std::sin(a), s->partials + 4 * kRedMaxBlocks);
  finish_kernel<1><<<1, kRedBlock, 0, s->stream>>>(s->partials + 4 * kRedMaxBlocks, nb, s->sums + kRbOut);
  HIPCHK(s, hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// Scratch of aero3d_readback_static: block partials of the residual and force
// sums (5 kRedMaxBlocks), of the Mach maximum and of the Trefftz sum (one
// kRedMaxBlocks each), the sums, and the segment record. Every sum is grid-
// strided over at most kRedMaxBlocks blocks, so the size does not depend on
// the field's size. Each device holds its own copy until the process ends;
// host code uses the address hipGetSymbolAddress returns for the current
// device, never the array's name.
// ---------------------------------------------------------------------------
constexpr size_t kScratchPart = 7 * (size_t)kRedMaxBlocks;        // doubles
constexpr size_t kScratchSums = 16;                                // doubles
constexpr size_t kScratchBytes = (kScratchPart + kScratchSums) * sizeof(double) + ((sizeof(RbSeg) + 255) / 256) * 256;
__device__ __align__(256) char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
constexpr int64_t kStaticMaxCells = int64_t(1) << 30;   // 32-bit element offsets in the kernels
constexpr int kStaticMaxForces = 1 << 30;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Address of this device's scratch, looked up once per device (the current one).
int scratch_of_device(void **out, std::string &err) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) { err = "readback_static: no usable HIP device"; return AERO3D_E_NODEV; }
  std::lock_guard<std::mutex> l(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) { err = std::string("readback_static: hipGetSymbolAddress: ") + hipGetErrorString(e); return AERO3D_E_HIP_BASE - (int)e; }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int aero3d_create(const aero3d_grid *g, const uint8_t *t, const float *phi, const float *ap0, const float *ap1,
                  const float *ap2, const float *ap3, aero3d_solver **out) {
  if (!out) return fail(nullptr, AERO3D_E_ARG, "aero3d_create: null output handle");
  *out = nullptr;
  if (!g || !g->fx || !g->fy || !g->fz || !t || !phi || !ap0 || !ap1 || !ap2 || !ap3)
    return fail(nullptr, AERO3D_E_ARG, "aero3d_create: null argument");
  // the y = 0 mirror ghosts of rows -1 and -2 are rows 0 and 1, so NY >= 2;
  // element offsets are formed in 32 bits, so N <= 2^30
  if (g->NX < 1 || g->NY < 2 || g->NZ < 1 || (int64_t)g->NX * g->NY * g->NZ > (int64_t)1 << 30)
    return fail(nullptr, AERO3D_E_ARG, "aero3d_create: grid size out of range (NX, NZ >= 1, NY >= 2, NX*NY*NZ <= 2^30)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(nullptr, AERO3D_E_NODEV, "aero3d_create: no HIP device");
  aero3d_solver *s = new aero3d_solver();
  s->NX = g->NX; s->NY = g->NY; s->NZ = g->NZ; s->N = g->NX * g->NY * g->NZ;
  // default stage-kernel shape: 32 x 8 tiles where 32-column tiles cover NX with
  // fewer padded columns than 64-column ones (the same bits either way), else 64 x 8
  s->tile_ty = (g->NX + 31) / 32 * 32 < (g->NX + 63) / 64 * 64 ? 32 : 8;
  s->ns = ((size_t)s->N + 63) / 64 * 64;
  for (int k = 0; k < 3; k++) std::memcpy(s->fix[k], s->F, sizeof s->F);
  auto bail = [&](int rc) { g_create_error = s->err; aero3d_destroy(s); return rc; };
  int rc;
  if (hipGetDevice(&s->device) != hipSuccess) return bail(fail(s, AERO3D_E_NODEV, "hipGetDevice failed"));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(s, AERO3D_E_NODEV, "hipStreamCreate failed"));
  {
    int ncu = 0, o4 = 0, o8 = 0, o32 = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess || ncu <= 0)
      return bail(fail(s, AERO3D_E_NODEV, "hipDeviceGetAttribute(MultiprocessorCount) failed"));
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o4, tile_kernel_bal<64, 4, true, 2>, 256, 0) != hipSuccess) o4 = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o8, tile_kernel_bal<64, 8, true, 2>, 512, 0) != hipSuccess) o8 = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o32, tile_kernel_bal<32, 8, true, 2>, 256, 0) != hipSuccess) o32 = 1;
    s->resident4 = std::max(1, ncu * std::max(1, o4));
    s->resident8 = std::max(1, ncu * std::max(1, o8));
    s->resident32 = std::max(1, ncu * std::max(1, o32));
  }
  // grid: FP32 widths and centres from FP64 faces, as the reference stores them
  std::vector<float> hdx(s->NX), hdy(s->NY), hdz(s->NZ), hcen;
  s->hxc.resize(s->NX);
  for (int i = 0; i < s->NX; i++) { hdx[i] = (float)(g->fx[i + 1] - g->fx[i]); s->hxc[i] = (float)(0.5 * (g->fx[i] + g->fx[i + 1])); }
  for (int j = 0; j < s->NY; j++) hdy[j] = (float)(g->fy[j + 1] - g->fy[j]);
  for (int k = 0; k < s->NZ; k++) hdz[k] = (float)(g->fz[k + 1] - g->fz[k]);
  s->hdy = hdy; s->hdz = hdz;
  hcen = s->hxc;
  for (int j = 0; j < s->NY; j++) hcen.push_back((float)(0.5 * (g->fy[j] + g->fy[j + 1])));
  for (int k = 0; k < s->NZ; k++) hcen.push_back((float)(0.5 * (g->fz[k] + g->fz[k + 1])));
  const size_t N = (size_t)s->N;
  if ((rc = dalloc(s, &s->dx, s->NX)) || (rc = dalloc(s, &s->dy, s->NY)) || (rc = dalloc(s, &s->dz, s->NZ)) ||
      (rc = dalloc(s, &s->cen, hcen.size())) ||
      (rc = dalloc(s, &s->t, N)) || (rc = dalloc(s, &s->phi, N)) || (rc = dalloc(s, &s->ap, 4 * N)) ||
      (rc = dalloc(s, &s->state, 15 * s->ns)) || (rc = dalloc(s, &s->dt, N)) || (rc = dalloc(s, &s->rsnap, N)) ||
      (rc = dalloc(s, &s->partials, 5 * kRedMaxBlocks)) || (rc = dalloc(s, &s->sums, 16)) || (rc = dalloc(s, &s->dseg, 1)))
    return bail(rc);
  auto up = [&](void *d, const void *h, size_t b) { return hipMemcpy(d, h, b, hipMemcpyHostToDevice); };
  if (up(s->dx, hdx.data(), 4 * s->NX) || up(s->dy, hdy.data(), 4 * s->NY) || up(s->dz, hdz.data(), 4 * s->NZ) ||
      up(s->cen, hcen.data(), 4 * hcen.size()) ||
      up(s->t, t, N) || up(s->phi, phi, 4 * N) || up(s->ap, ap0, 4 * N) || up(s->ap + N, ap1, 4 * N) ||
      up(s->ap + 2 * N, ap2, 4 * N) || up(s->ap + 3 * N, ap3, 4 * N) ||
      hipMemset(s->state, 0, sizeof(float) * 15 * s->ns) || hipMemset(s->dt, 0, 4 * N))
    return bail(fail(s, AERO3D_E_HIP_BASE, "aero3d_create: upload failed"));
  *out = s;
  return 0;
}

int aero3d_destroy(aero3d_solver *s) {
  if (!s) return 0;
  // teardown: an error here leaves nothing to recover, so return codes are not checked
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  void *ptrs[] = {s->dx, s->dy, s->dz, s->cen, s->t, s->phi, s->ap, s->state, s->dt, s->rsnap, s->fq, s->fcx, s->fcz, s->fx, s->fz,
                  s->partials, s->sums, s->dseg, s->probe};
  for (void *p : ptrs) if (p) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

const char *aero3d_last_error(const aero3d_solver *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

int aero3d_set_free(aero3d_solver *s, const double F[5]) {
  if (!s || !F) return fail(s, AERO3D_E_ARG, "set_free: null argument");
  std::memcpy(s->F, F, sizeof s->F);
  return 0;
}
int aero3d_set_fix(aero3d_solver *s, const double fix[15]) {
  if (!s || !fix) return fail(s, AERO3D_E_ARG, "set_fix: null argument");
  std::memcpy(s->fix, fix, sizeof s->fix);
  return apply_fix(s);
}
int aero3d_set_opt(aero3d_solver *s, double cfl, int order) {
  if (!s) return AERO3D_E_ARG;
  if (!(cfl > 0) || !(cfl < 10) || (order != 1 && order != 2)) return fail(s, AERO3D_E_ARG, "set_opt: cfl must be in (0,10), order 1 or 2");
  s->cfl = cfl; s->order = order;
  return 0;
}
int aero3d_reset(aero3d_solver *s) {
  if (!s) return AERO3D_E_ARG;
  s->cur = 0; s->it = 0; s->snap_it = -1;
  fill_kernel<<<(s->N + kBlock - 1) / kBlock, kBlock, 0, s->stream>>>(buf(s, 0), s->ns, s->N, (float)s->F[0], (float)s->F[1],
                                                                      (float)s->F[2], (float)s->F[3], (float)s->F[4]);
  HIPCHK(s, hipGetLastError());
  return apply_fix(s);
}
int aero3d_step(aero3d_solver *s, int n) {
  if (!s || n < 0) return fail(s, AERO3D_E_ARG, "step: n must be >= 0");
  return enqueue_steps(s, n);
}
int aero3d_iter(const aero3d_solver *s, int *it) {
  if (!s || !it) return AERO3D_E_ARG;
  *it = s->it;
  return 0;
}
int aero3d_sync(aero3d_solver *s) {
  if (!s) return AERO3D_E_ARG;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return 0;
}
int aero3d_read(aero3d_solver *s, float *r, float *ru, float *rv, float *rw, float *E, float *dt) {
  if (!s) return AERO3D_E_ARG;
  float *Q = buf(s, s->cur);
  float *dst[5] = {r, ru, rv, rw, E};
  const size_t b = sizeof(float) * (size_t)s->N;
  for (int c = 0; c < 5; c++)
    if (dst[c]) HIPCHK(s, hipMemcpyAsync(dst[c], Q + c * s->ns, b, hipMemcpyDeviceToHost, s->stream));
  if (dt) HIPCHK(s, hipMemcpyAsync(dt, s->dt, b, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return 0;
}
int aero3d_write(aero3d_solver *s, const float *r, const float *ru, const float *rv, const float *rw, const float *E) {
  if (!s || !r || !ru || !rv || !rw || !E) return fail(s, AERO3D_E_ARG, "write: null argument");
  float *Q = buf(s, s->cur);
  const float *src[5] = {r, ru, rv, rw, E};
  for (int c = 0; c < 5; c++)
    HIPCHK(s, hipMemcpyAsync(Q + c * s->ns, src[c], sizeof(float) * (size_t)s->N, hipMemcpyHostToDevice, s->stream));
  if (int rc = apply_fix(s)) return rc;
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return 0;
}

int aero3d_set_force_list(aero3d_solver *s, const int32_t *q, const double *cx, const double *cz, const double *x,
                          const double *z, int n) {
  if (!s || n < 0 || (n > 0 && (!q || !cx || !cz || !x || !z))) return fail(s, AERO3D_E_ARG, "set_force_list: bad argument");
  for (int k = 0; k < n; k++)
    if (q[k] < 0 || q[k] >= s->N) return fail(s, AERO3D_E_ARG, "set_force_list: cell index out of range");
  HIPCHK(s, hipStreamSynchronize(s->stream));
  void *old[] = {s->fq, s->fcx, s->fcz, s->fx, s->fz};
  hipError_t fe = hipSuccess;
  for (void *p : old)
    if (p) { const hipError_t e = hipFree(p); if (fe == hipSuccess) fe = e; }
  s->fq = nullptr; s->fcx = s->fcz = s->fx = s->fz = nullptr;
  s->nf = 0;                 // no list until the new one is in place
  HIPCHK(s, fe);
  int rc;
  if ((rc = dalloc(s, &s->fq, n)) || (rc = dalloc(s, &s->fcx, n)) || (rc = dalloc(s, &s->fcz, n)) ||
      (rc = dalloc(s, &s->fx, n)) || (rc = dalloc(s, &s->fz, n)))
    return rc;
  if (n > 0) {
    HIPCHK(s, hipMemcpy(s->fq, q, 4 * (size_t)n, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->fcx, cx, 8 * (size_t)n, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->fcz, cz, 8 * (size_t)n, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->fx, x, 8 * (size_t)n, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->fz, z, 8 * (size_t)n, hipMemcpyHostToDevice));
  }
  s->nf = n;
  return 0;
}

int aero3d_forces(aero3d_solver *s, double alphaDeg, double M, double S, double xcg, double zcg, double cref, double out[3]) {
  if (!s || !out) return AERO3D_E_ARG;
  if (!s->fq) return fail(s, AERO3D_E_STATE, "forces: no force list set");
  double o[kRbOut];
  if (int rc = readback_sums(s, false, true, alphaDeg, M, S, xcg, zcg, cref, o)) return rc;
  out[0] = o[6]; out[1] = o[7]; out[2] = o[8];
  return 0;
}

int aero3d_trefftz(aero3d_solver *s, double xT, double alphaDeg, double M, double S, double *cdi) {
  if (!s || !cdi) return AERO3D_E_ARG;
  if (s->NX < 2) return fail(s, AERO3D_E_ARG, "trefftz: needs NX >= 2");
  if (int rc = enqueue_trefftz(s, xT, alphaDeg)) return rc;
  double sum;
  HIPCHK(s, hipMemcpyAsync(&sum, s->sums + kRbOut, sizeof sum, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  const double Mn = std::max(M, 0.05);
  *cdi = 2 * sum / (Mn * Mn * S);
  return 0;
}

int aero3d_residual(aero3d_solver *s, double *res) {
  if (!s || !res) return AERO3D_E_ARG;
  double o[kRbOut];
  if (int rc = readback_sums(s, true, false, 0, 1, 1, 0, 0, 1, o)) return rc;
  *res = o[5];
  return 0;
}

int aero3d_readback(aero3d_solver *s, const aero3d_readback_in *in, aero3d_readback_out *out) {
  if (!s || !in || !out) return fail(s, AERO3D_E_ARG, "readback: null argument");
  const bool tz = !std::isnan(in->xT);
  if (tz && s->NX < 2) return fail(s, AERO3D_E_ARG, "readback: the Trefftz plane needs NX >= 2");
  const double Mn = std::max(in->M, 0.05);     // the readback's Mach floor (src/air_cfd.js:558, 561)
  const int it = s->it;
  // enqueued: Trefftz plane, Mach maximum, then the residual and forces (the residual retakes the snapshot)
  if (tz) { if (int rc = enqueue_trefftz(s, in->xT, in->alpha)) return rc; }
  const int nbm = reduce_blocks((s->N + 1) / 2);
  machmax_kernel<<<nbm, kRedBlock, 0, s->stream>>>(buf(s, s->cur), s->ns, s->t, s->N, s->partials + 3 * kRedMaxBlocks);
  machmax_finish<<<1, 64, 0, s->stream>>>(s->partials + 3 * kRedMaxBlocks, nbm, s->sums + kRbOut + 1);
  HIPCHK(s, hipGetLastError());
  double o[kRbOut + 2];
  if (int rc = readback_sums(s, true, s->fq != nullptr, in->alpha, Mn, in->S, in->xcg, in->zcg, in->cref, o, true)) return rc;
  out->it = it;
  out->res = o[5];
  out->CL = s->fq ? o[6] : NAN; out->CD = s->fq ? o[7] : NAN; out->Cm = s->fq ? o[8] : NAN;
  out->CDi = tz ? 2 * o[kRbOut] / (Mn * Mn * in->S) : NAN;
  out->Mmax = std::sqrt(o[kRbOut + 1]);
  return 0;
}

int aero3d_probe(aero3d_solver *s, int n, const double *xyz, const aero3d_probe_fs *fs, double M, const aero3d_probe_out *out) {
  if (!s || n < 0 || (n > 0 && !xyz) || !out) return fail(s, AERO3D_E_ARG, "probe: null argument or n < 0");
  if (n > kProbeMaxPoints) return fail(s, AERO3D_E_ARG, "probe: more than 2^26 points");
  if (out->state && !fs) return fail(s, AERO3D_E_ARG, "probe: the display state needs the freestream record");
  if (s->NX < 2 || s->NY < 2 || s->NZ < 2) return fail(s, AERO3D_E_ARG, "probe: needs at least 2 cells on each axis");
  if (n == 0) return 0;
  // one allocation: points, then each result array, 256-byte aligned
  const size_t N = (size_t)n;
  const size_t sz[8] = {24 * N, 32 * N, 32 * N, N, 40 * N, 96 * N, 24 * N, N};
  size_t off[8], total = 0;
  for (int c = 0; c < 8; c++) { off[c] = total; total += (sz[c] + 255) & ~(size_t)255; }
  if (total > s->probe_cap) {
    HIPCHK(s, hipStreamSynchronize(s->stream));
    if (s->probe) { HIPCHK(s, hipFree(s->probe)); s->bytes -= (int64_t)s->probe_cap; s->probe = nullptr; s->probe_cap = 0; }
    HIPCHK(s, hipMalloc((void **)&s->probe, total));
    s->probe_cap = total;
    s->bytes += (int64_t)total;
  }
  char *b = s->probe;
  ProbeArgs a{};
  a.NX = s->NX; a.NY = s->NY; a.NZ = s->NZ; a.n = n;
  a.xc = s->cen; a.yc = s->cen + s->NX; a.zc = s->cen + s->NX + s->NY;
  a.t = s->t; a.Q = buf(s, s->cur); a.ns = s->ns;
  a.pts = (const double *)(b + off[0]);
  if (fs) { a.fsT = fs->T; a.fsa = fs->a; a.fsP = fs->P; a.fsrho = fs->rho; a.fsq = fs->q; }
  a.M = M;
  a.idx = out->idx ? (int32_t *)(b + off[1]) : nullptr;
  a.w = out->w ? (float *)(b + off[2]) : nullptr;
  a.ok = out->ok ? (uint8_t *)(b + off[3]) : nullptr;
  a.W = out->W ? (double *)(b + off[4]) : nullptr;
  a.st = out->state ? (double *)(b + off[5]) : nullptr;
  a.vel = out->vel ? (double *)(b + off[6]) : nullptr;
  a.vok = out->vok ? (uint8_t *)(b + off[7]) : nullptr;
  HIPCHK(s, hipMemcpyAsync(b + off[0], xyz, sz[0], hipMemcpyHostToDevice, s->stream));
  probe_kernel<<<(n + 255) / 256, 256, 0, s->stream>>>(a);
  HIPCHK(s, hipGetLastError());
  void *dst[8] = {nullptr, out->idx, out->w, out->ok, out->W, out->state, out->vel, out->vok};
  for (int c = 1; c < 8; c++)
    if (dst[c]) HIPCHK(s, hipMemcpyAsync(dst[c], b + off[c], sz[c], hipMemcpyDeviceToHost, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  return 0;
}

int aero3d_readback_static(const aero3d_field *f, const aero3d_readback_in *in, aero3d_readback_out *out, void *stream) {
  static thread_local std::string err;
  if (!f || !in || !out) { err = "readback_static: null argument"; g_create_error = err; return AERO3D_E_ARG; }
  const int64_t N = (int64_t)f->NX * f->NY * f->NZ;
  if (f->NX < 1 || f->NY < 1 || f->NZ < 1 || N > kStaticMaxCells || f->ns < (uint64_t)N || f->nf < 0 || f->nf > kStaticMaxForces ||
      f->nit < 0 || !f->Q || !f->t || !f->rsnap || (f->nf > 0 && (!f->fq || !f->cx || !f->cz || !f->x || !f->z))) {
    g_create_error = "readback_static: bad field (NX, NY, NZ >= 1, NX*NY*NZ <= 2^30, ns >= N, 0 <= nf <= 2^30, nit >= 0, non-null arrays)";
    return AERO3D_E_ARG;
  }
  const bool tz = !std::isnan(in->xT);
  if (tz && (f->NX < 2 || !f->xc || !f->dy || !f->dz)) { g_create_error = "readback_static: the Trefftz plane needs NX >= 2, xc, dy and dz"; return AERO3D_E_ARG; }
  void *base = nullptr;
  if (int rc = scratch_of_device(&base, err)) { g_create_error = err; return rc; }
  double *part = (double *)base, *sums = part + kScratchPart;
  RbSeg *dseg = (RbSeg *)(sums + kScratchSums);
  hipStream_t st = (hipStream_t)stream;
  const double Mn = std::max(in->M, 0.05);
  auto chk = [&](hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    g_create_error = std::string("readback_static: ") + what + ": " + hipGetErrorString(e);
    return AERO3D_E_HIP_BASE - (int)e;
  };
  int rc;
  const size_t ns = (size_t)f->ns;
  if (tz) {   // cfdLocate on the caller's FP32 centres (src/air_cfd.js:453), then the plane sum
    const float *c = f->xc;
    const int n = f->NX;
    int i0; double tx;
    if (in->xT <= c[0]) { i0 = 0; tx = 0; }
    else if (in->xT >= c[n - 1]) { i0 = n - 2; tx = 1; }
    else {
      int lo = 0, hi = n - 1;
      while (hi - lo > 1) { const int m = (lo + hi) >> 1; if (c[m] <= in->xT) lo = m; else hi = m; }
      i0 = lo; tx = (in->xT - c[lo]) / ((double)c[hi] - c[lo]);
    }
    const double a = in->alpha * (M_PI / 180);
    const int nb = reduce_blocks(f->NY * f->NZ);
    trefftz_kernel<<<nb, kRedBlock, 0, st>>>(f->Q, ns, f->t, f->dy, f->dz, f->NX, f->NY, f->NZ, i0, tx, std::cos(a), std::sin(a),
                                            part + 6 * kRedMaxBlocks);
    finish_kernel<1><<<1, kRedBlock, 0, st>>>(part + 6 * kRedMaxBlocks, nb, sums + kRbOut);
  }
  const int nbm = reduce_blocks((int)((N + 1) / 2));
  machmax_kernel<<<nbm, kRedBlock, 0, st>>>(f->Q, ns, f->t, (int)N, part + 5 * kRedMaxBlocks);
  machmax_finish<<<1, 64, 0, st>>>(part + 5 * kRedMaxBlocks, nbm, sums + kRbOut + 1);
  RbSeg g{};
  g.Q = f->Q; g.ns = ns; g.t = f->t; g.N = (int)N; g.rsnap = f->rsnap; g.nit = f->nit;
  g.fq = f->fq; g.cx = f->cx; g.cz = f->cz; g.x = f->x; g.z = f->z; g.nf = f->nf;
  g.xcg = in->xcg; g.zcg = in->zcg;
  const double a = in->alpha * (M_PI / 180);
  g.ca = std::cos(a); g.sa = std::sin(a); g.kq = 2 / (0.5 * Mn * Mn * in->S); g.cref = in->cref;
  g.nb_res = reduce_blocks((int)((N + 2) / 3));
  g.nb_f = f->nf > 0 ? reduce_blocks(f->nf) : 0;
  g.part = part; g.out = sums;
  if ((rc = chk(hipMemcpyAsync(dseg, &g, sizeof g, hipMemcpyHostToDevice, st), "segment copy"))) return rc;
  rb_partials_kernel<<<g.nb_res + g.nb_f, kRedBlock, 0, st>>>(dseg, 1);
  rb_finish_kernel<<<1, kRedBlock, 0, st>>>(dseg);
  if ((rc = chk(hipGetLastError(), "launch"))) return rc;
  double o[kRbOut + 2];
  if ((rc = chk(hipMemcpyAsync(o, sums, sizeof o, hipMemcpyDeviceToHost, st), "result copy"))) return rc;
  if ((rc = chk(hipStreamSynchronize(st), "wait"))) return rc;
  out->it = -1;
  out->res = o[5];
  out->CL = f->nf > 0 ? o[6] : NAN; out->CD = f->nf > 0 ? o[7] : NAN; out->Cm = f->nf > 0 ? o[8] : NAN;
  out->CDi = tz ? 2 * o[kRbOut] / (Mn * Mn * in->S) : NAN;
  out->Mmax = std::sqrt(o[kRbOut + 1]);
  return 0;
}

int aero3d_time_steps(aero3d_solver *s, int n, float *ms) {
  if (!s || !ms || n < 0) return AERO3D_E_ARG;
  struct Ev { hipEvent_t e = nullptr; ~Ev() { if (e) (void)hipEventDestroy(e); } } e0, e1;   // released on every path
  HIPCHK(s, hipEventCreate(&e0.e));
  HIPCHK(s, hipEventCreate(&e1.e));
  HIPCHK(s, hipEventRecord(e0.e, s->stream));
  int rc = enqueue_steps(s, n);
  if (rc) return rc;
  HIPCHK(s, hipEventRecord(e1.e, s->stream));
  HIPCHK(s, hipEventSynchronize(e1.e));
  HIPCHK(s, hipEventElapsedTime(ms, e0.e, e1.e));
  return 0;
}

int aero3d_set_tile(aero3d_solver *s, int ty, int zchunk) {
  if (!s) return AERO3D_E_ARG;
  if (!(ty == 0 || ty == 4 || ty == 8 || ty == 32) || zchunk < -8)
    return fail(s, AERO3D_E_ARG, "set_tile: ty must be 0, 4, 8 or 32 and zchunk >= -8");
  s->tile_ty = ty;
  s->tile_zc = zchunk;
  return 0;
}
int aero3d_get_tile(const aero3d_solver *s, int *ty, int *zchunk) {
  if (!s || !ty || !zchunk) return AERO3D_E_ARG;
  *ty = s->tile_ty; *zchunk = s->tile_zc;
  return 0;
}

int aero3d_autotune(aero3d_solver *s, int iters, int *ty, int *zchunk, float *ms_best) {
  if (!s || iters < 1) return AERO3D_E_ARG;
  // candidates: both tile heights, z chunks from short (many blocks) to long (little preamble)
  const int tys[] = {4, 8, 32};
  const int zcs[] = {0, -2, -3, 2, 3, 4, 6, 8};
  // time on a scratch copy of the state: save the current state and the
  // local dt of the last iteration, restore both after
  const size_t b = sizeof(float) * 5 * s->ns, bdt = sizeof(float) * (size_t)s->N;
  float *save = nullptr;
  HIPCHK(s, hipMalloc((void **)&save, b + bdt));
  HIPCHK(s, hipMemcpyAsync(save, buf(s, s->cur), b, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(s, hipMemcpyAsync(save + 5 * s->ns, s->dt, bdt, hipMemcpyDeviceToDevice, s->stream));
  const int cur0 = s->cur, it0 = s->it;
  int bty = s->tile_ty, bzc = s->tile_zc;
  float best = 1e30f;
  for (int t : tys)
    for (int z : zcs) {
      if (z > s->NZ) continue;
      s->tile_ty = t; s->tile_zc = z;
      float ms;
      int rc = aero3d_time_steps(s, 1, &ms);          // warm
      if (!rc) rc = aero3d_time_steps(s, iters, &ms);
      if (rc) { (void)hipFree(save); return rc; }
      if (ms < best) { best = ms; bty = t; bzc = z; }
    }
  s->tile_ty = bty; s->tile_zc = bzc;
  s->cur = cur0; s->it = it0;
  HIPCHK(s, hipMemcpyAsync(buf(s, s->cur), save, b, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(s, hipMemcpyAsync(s->dt, save + 5 * s->ns, bdt, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(s, hipStreamSynchronize(s->stream));
  HIPCHK(s, hipFree(save));
  if (ty) *ty = bty;
  if (zchunk) *zchunk = bzc;
  if (ms_best) *ms_best = best / iters;
  return 0;
}

int aero3d_footprint(const aero3d_solver *s, int64_t *bytes) {
  if (!s || !bytes) return AERO3D_E_ARG;
  *bytes = s->bytes;
  return 0;
}

}  // extern "C"
