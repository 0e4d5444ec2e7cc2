// ============================================================================
// aero3d_batch.hip — batched and served forms of the 3D Euler cut-cell solver.
//
// Batch: many independent solves of uneven grid sizes stepped together.
//   * one device arena (single hipMalloc) holds every registered geometry
//     (grid widths, flags, phi, apertures, force list) and every case's
//     state; a host-side best-fit free list hands out aligned ranges, so
//     admitting and retiring cases never calls hipMalloc / hipFree;
//   * a device case table lists, per active case, its CaseView, stage
//     constants, state buffers and the range of work items (64 x TY x zc
//     tile chunks) it owns;
//   * one launch per RK stage covers all work items of all cases of a group:
//     by default one block per item (grid = the group's item count, XCD-
//     grouped so neighbouring tiles share an L2), or, with
//     aero3d_batch_set_persistent, the device's resident blocks walking the
//     items grid-stride;
//   * three SSP-RK2 iterations (six launches; the three-buffer rotation has
//     period three) are captured in a hipGraph and replayed; the table lives
//     in device memory, so a changed table re-captures a group's graph only
//     when its launch grid (item count) changed, and never in the persistent
//     form.
// Server: a FIFO of solve requests admitted by HBM footprint into the batch
// as memory frees up (continuous batching); each tick steps every resident
// case by the check interval, measures each case's density residual on the
// device, retires converged / capped cases with their force coefficients.
// ============================================================================
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "aero3d.h"
#include "euler3d.hpp"
#include "euler3d_tile.hpp"
#include "reduce.hpp"

using namespace aero3d;

namespace {

constexpr int kTY = 8;              // tile height of the batched kernel (64 x 8 columns)
constexpr size_t kAlign = 256;

struct SlotDev {                    // one case in the device case table
  CaseView cv;
  StageConsts kc;
  float *state;                     // 3 buffers x 5 planes x ns
  unsigned long long ns;
  int cur;                          // buffer holding Q at phase 0
  int ntx, nty, zc, item0, nitems;
};
struct TableHdr { int nslots, total; };

// Stage kernel over the case table; a work item is (case, tile column, z
// chunk). persist = 0: one item per block, the launch grid is the group's item
// count, so each XCD takes a contiguous range of items. persist = 1: the
// resident blocks walk all items (one persistent launch whatever the table);
// the loop state stays live across the tile stage and spills to scratch, so on
// 16 cases of the batched workload it runs 11.2 against 13.1 G cell-updates/s.
// phase: iterations since the table's buffer indices were written (mod 3);
// stage: 0 or 1 of SSP-RK2. At least four waves per SIMD (<= 128 VGPRs).
template <int stage, int persist>
__global__ void __launch_bounds__(64 * kTY) __attribute__((amdgpu_waves_per_eu(4))) batch_kernel(const TableHdr *__restrict__ hdr, const SlotDev *__restrict__ slots,
                                                          int phase) {
  __shared__ TileSmem<64, kTY> sm;
  const int total = hdr->total, nsl = hdr->nslots;
  const int G = gridDim.x, r = xcd_remap(blockIdx.x, G);
  auto run = [&](int item) {
    int lo = 0, hi = nsl - 1;
    while (lo < hi) { const int m = (lo + hi + 1) >> 1; if (slots[m].item0 <= item) lo = m; else hi = m - 1; }
    lo = __builtin_amdgcn_readfirstlane(lo);      // block-uniform: keep the case record in scalar registers
    const SlotDev &S = slots[lo];
    const int li = item - S.item0;
    const int bx = li % S.ntx, by = (li / S.ntx) % S.nty, bz = li / (S.ntx * S.nty);
    const int k0 = bz * S.zc, k1 = min(S.cv.NZ, k0 + S.zc);
    const int q = (S.cur + 2 * phase) % 3;
    const size_t plane5 = 5 * (size_t)S.ns;
    float *Q = S.state + (size_t)q * plane5, *S1 = S.state + (size_t)((q + 1) % 3) * plane5,
          *Bb = S.state + (size_t)((q + 2) % 3) * plane5;
    if constexpr (stage == 0) tile_stage<64, kTY, false, 1>(S.cv, S.kc, Q, Q, S1, nullptr, S.ns, 0.f, 1.f, bx * 64, by * kTY, k0, k1, sm);
    else tile_stage<64, kTY, false, 2>(S.cv, S.kc, S1, Q, Bb, nullptr, S.ns, 0.5f, 0.5f, bx * 64, by * kTY, k0, k1, sm);
  };
  if constexpr (persist) {
    for (int base = 0; base < total; base += G) {
      if (base > 0) __syncthreads();              // the next item reuses the LDS tile
      if (base + r < total) run(base + r);
    }
  } else {
    if (r < total) run(r);
  }
}

__global__ void bfill_kernel(float *__restrict__ dst, size_t ns, int N, float f0, float f1, float f2, float f3, float f4) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= N) return;
  dst[q] = f0; dst[ns + q] = f1; dst[2 * ns + q] = f2; dst[3 * ns + q] = f3; dst[4 * ns + q] = f4;
}
// Fixed cells (flags 3..5) := their conserved fixed states (the stage kernels
// take a fixed cell's primitives from the stored state).
struct BFixStates { float q[3][5]; };
__global__ void bfixcells_kernel(float *__restrict__ dst, size_t ns, int N, const uint8_t *__restrict__ t, BFixStates f) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= N) return;
  const int g = t[q] - 3;
  if (g < 0) return;
  for (int c = 0; c < 5; c++) dst[c * ns + q] = f.q[g][c];
}
__global__ void bcopy_kernel(float *__restrict__ dst, const float *__restrict__ src, int N) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < N) dst[q] = src[q];
}

// ---------------------------------------------------------------------------
// device arena with a host-side best-fit free list
// ---------------------------------------------------------------------------
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  std::map<size_t, size_t> free_;          // offset -> size
  bool init(size_t bytes) {
    if (hipMalloc((void **)&base, bytes) != hipSuccess) return false;
    cap = bytes; free_.clear(); free_[0] = bytes; used = 0;
    return true;
  }
  void *get(size_t bytes) {
    bytes = (bytes + kAlign - 1) / kAlign * kAlign;
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes && (best == free_.end() || it->second < best->second)) best = it;
    if (best == free_.end()) return nullptr;
    const size_t off = best->first, sz = best->second;
    free_.erase(best);
    if (sz > bytes) free_[off + bytes] = sz - bytes;
    used += bytes;
    return base + off;
  }
  void put(void *p, size_t bytes) {
    if (!p) return;
    bytes = (bytes + kAlign - 1) / kAlign * kAlign;
    size_t off = (char *)p - base;
    used -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) { it->second += nx->second; free_.erase(nx); }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) { pv->second += it->second; free_.erase(it); }
    }
  }
};

struct Geometry {
  int NX, NY, NZ, N;
  void *mem = nullptr; size_t bytes = 0;
  float *dx, *dy, *dz; uint8_t *t; float *phi, *ap;
  int nf = 0; int32_t *fq = nullptr; double *fcx = nullptr, *fcz = nullptr, *fx = nullptr, *fz = nullptr;
  std::vector<float> hdy, hdz, hxc;
  int refs = 0;
  bool dead = false;
};

struct Slot {
  bool used = false;
  int geo = -1;
  void *mem = nullptr; size_t bytes = 0;
  float *state = nullptr, *rsnap = nullptr;
  double *partials = nullptr;       // FP64 block partials and the slot's sums
  size_t ns = 0;
  int cur = 0, it = 0, snap_it = -1;
  StageConsts kc;
  int zc = 3;
};

}  // namespace

struct aero3d_batch {
  Arena arena;
  hipStream_t stream = nullptr;
  std::vector<Geometry> geos;
  std::vector<Slot> slots;
  TableHdr *d_hdr = nullptr;         // one header per case group
  SlotDev *d_table = nullptr;        // max_slots entries per case group
  int max_slots = 0;
  int grid = 0;                      // resident blocks of the persistent form
  std::vector<hipGraph_t> graphs;    // per case group: three iterations
  std::vector<hipGraphExec_t> execs;
  std::vector<int> totals;           // per case group: work items (the launch grid)
  std::vector<int> graph_total;      // per case group: the launch grid its graph was captured with
  int persistent = 0;                // 1: resident blocks walk the items (batch_kernel<., 1>)
  bool table_dirty = true;
  std::vector<int> order;            // slot ids in table order
  int ngroups = 0;
  int64_t group_cells = int64_t(1) << 40;  // cases are stepped in groups of about this many cells (default: one group)
  int zc = 3;
  std::string err;
};

namespace {

int bfail(aero3d_batch *b, int code, const std::string &m) { if (b) b->err = m; return code; }
#define BCHK(b, expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return bfail((b), AERO3D_E_HIP_BASE - (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

size_t slot_bytes(const Geometry &g, size_t *ns_out) {
  const size_t ns = ((size_t)g.N + 63) / 64 * 64;
  if (ns_out) *ns_out = ns;
  auto al = [](size_t b) { return (b + kAlign - 1) / kAlign * kAlign; };
  return al(4 * 15 * ns) + al(4 * (size_t)g.N) + al(8 * (3 * kRedMaxBlocks + 8));
}

void prim5h(const double *a, double *o) {
  const double r = std::max(a[0], 1e-6), u = a[1] / r, v = a[2] / r, w = a[3] / r;
  o[0] = r; o[1] = u; o[2] = v; o[3] = w;
  o[4] = std::max((1.4 - 1) * (a[4] - 0.5 * r * (u * u + v * v + w * w)), 1e-7);
}

// Rewrite the device case tables from the active slots. Cases are split into
// groups of about group_cells cells (a group's state and masks stay in the
// Infinity Cache across its stages and iterations); each group has its own
// header and table region, stepped by its own launches.
int upload_table(aero3d_batch *b) {
  b->order.clear();
  std::vector<std::vector<SlotDev>> groups(1);
  int64_t cells = 0;
  for (int s = 0; s < (int)b->slots.size(); s++) {
    Slot &S = b->slots[s];
    if (!S.used) continue;
    const Geometry &g = b->geos[S.geo];
    if (!groups.back().empty() && cells + g.N > b->group_cells) { groups.emplace_back(); cells = 0; }
    cells += g.N;
    SlotDev d;
    d.cv = CaseView{g.NX, g.NY, g.NZ, g.N, g.dx, g.dy, g.dz, g.t, g.phi, g.ap};
    d.kc = S.kc;
    d.state = S.state; d.ns = S.ns; d.cur = S.cur;
    d.zc = S.zc;
    d.ntx = (g.NX + 63) / 64; d.nty = (g.NY + kTY - 1) / kTY;
    d.nitems = d.ntx * d.nty * ((g.NZ + S.zc - 1) / S.zc);
    d.item0 = 0;
    groups.back().push_back(d);
    b->order.push_back(s);
  }
  if (groups.back().empty()) groups.pop_back();
  b->ngroups = (int)groups.size();
  std::vector<TableHdr> hdrs(groups.size());
  for (size_t gi = 0; gi < groups.size(); gi++) {
    int item = 0;
    for (auto &d : groups[gi]) { d.item0 = item; item += d.nitems; }
    hdrs[gi] = TableHdr{(int)groups[gi].size(), item};
    BCHK(b, hipMemcpyAsync(b->d_table + gi * (size_t)b->max_slots, groups[gi].data(), sizeof(SlotDev) * groups[gi].size(),
                           hipMemcpyHostToDevice, b->stream));
  }
  if (!hdrs.empty()) BCHK(b, hipMemcpyAsync(b->d_hdr, hdrs.data(), sizeof(TableHdr) * hdrs.size(), hipMemcpyHostToDevice, b->stream));
  b->totals.clear();
  for (auto &h : hdrs) b->totals.push_back(h.total);
  // the host copies above must stay alive until the copies run: wait here (table changes are rare)
  BCHK(b, hipStreamSynchronize(b->stream));
  b->table_dirty = false;
  return 0;
}

// Both stages of one SSP-RK2 iteration of case group gi at rotation phase ph.
void launch_iteration(aero3d_batch *b, int gi, int ph) {
  const TableHdr *h = b->d_hdr + gi;
  const SlotDev *t = b->d_table + gi * (size_t)b->max_slots;
  if (b->persistent) {
    batch_kernel<0, 1><<<b->grid, dim3(64, kTY), 0, b->stream>>>(h, t, ph);
    batch_kernel<1, 1><<<b->grid, dim3(64, kTY), 0, b->stream>>>(h, t, ph);
  } else {
    const int G = std::max(1, b->totals[gi]);
    batch_kernel<0, 0><<<G, dim3(64, kTY), 0, b->stream>>>(h, t, ph);
    batch_kernel<1, 0><<<G, dim3(64, kTY), 0, b->stream>>>(h, t, ph);
  }
}

// the launch grid a group's iteration uses (what its captured graph holds)
int group_grid(const aero3d_batch *b, int gi) { return b->persistent ? -b->grid : std::max(1, b->totals[gi]); }

int build_graph(aero3d_batch *b, int gi) {
  hipGraph_t g;
  hipGraphExec_t e;
  BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int ph = 0; ph < 3; ph++) launch_iteration(b, gi, ph);
  BCHK(b, hipStreamEndCapture(b->stream, &g));
  BCHK(b, hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
  b->graphs[gi] = g; b->execs[gi] = e; b->graph_total[gi] = group_grid(b, gi);
  return 0;
}

int batch_steps(aero3d_batch *b, int n) {
  int rc;
  if (b->table_dirty && (rc = upload_table(b))) return rc;
  if (b->order.empty() || n <= 0) return 0;
  if ((int)b->execs.size() < b->ngroups) {
    b->execs.resize(b->ngroups, nullptr); b->graphs.resize(b->ngroups, nullptr); b->graph_total.resize(b->ngroups, -1);
  }
  const int r = n % 3;
  for (int gi = 0; gi < b->ngroups; gi++) {
    // a graph holds its launch grid: capture again when it changed
    if (b->execs[gi] && b->graph_total[gi] != group_grid(b, gi)) {
      (void)hipGraphExecDestroy(b->execs[gi]); (void)hipGraphDestroy(b->graphs[gi]);
      b->execs[gi] = nullptr; b->graphs[gi] = nullptr;
    }
    if (n >= 3 && !b->execs[gi] && (rc = build_graph(b, gi))) return rc;
    for (int k = 0; k + 3 <= n; k += 3) BCHK(b, hipGraphLaunch(b->execs[gi], b->stream));
    for (int ph = 0; ph < r; ph++) launch_iteration(b, gi, ph);
  }
  BCHK(b, hipGetLastError());
  for (int s : b->order) { b->slots[s].cur = (b->slots[s].cur + 2 * r) % 3; b->slots[s].it += n; }
  if (r) b->table_dirty = true;      // buffer indices moved
  return 0;
}

inline float *slot_Q(const Slot &S) { return S.state + (size_t)S.cur * 5 * S.ns; }

// Enqueue the residual of slot S into its sums area (sums[0..1]); snapshot after.
void enqueue_residual(aero3d_batch *b, Slot &S) {
  const Geometry &g = b->geos[S.geo];
  double *partials = S.partials, *sums = S.partials + 3 * kRedMaxBlocks;
  const float *r = slot_Q(S);
  if (S.snap_it >= 0 && S.it > S.snap_it) {
    const int nb = reduce_blocks((g.N + 2) / 3);
    residual_kernel<<<nb, kRedBlock, 0, b->stream>>>(r, S.rsnap, g.t, g.N, partials);
    finish_kernel<2><<<1, kRedBlock, 0, b->stream>>>(partials, nb, sums);
  }
  bcopy_kernel<<<(g.N + 255) / 256, 256, 0, b->stream>>>(S.rsnap, r, g.N);
}

void enqueue_forces(aero3d_batch *b, Slot &S, double xcg, double zcg) {
  const Geometry &g = b->geos[S.geo];
  double *partials = S.partials, *sums = S.partials + 3 * kRedMaxBlocks + 2;
  const int nb = reduce_blocks(g.nf);
  forces_kernel<<<nb, kRedBlock, 0, b->stream>>>(slot_Q(S), S.ns, g.fq, g.fcx, g.fcz, g.fx, g.fz, g.nf, xcg, zcg, partials);
  finish_kernel<3><<<1, kRedBlock, 0, b->stream>>>(partials, nb, sums);
}

void coefficients(const double *F3, double alphaDeg, double M, double S, double cref, double *out) {
  const double a = alphaDeg * (M_PI / 180), q = 0.5 * M * M, k = 2 / (q * S);
  out[0] = k * (F3[1] * std::cos(a) - F3[0] * std::sin(a));
  out[1] = k * (F3[0] * std::cos(a) + F3[1] * std::sin(a));
  out[2] = k * F3[2] / cref;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI: batch
// ---------------------------------------------------------------------------
extern "C" {

int aero3d_batch_create(int64_t arena_bytes, int max_slots, aero3d_batch **out) {
  if (!out || arena_bytes <= 0 || max_slots < 1) return AERO3D_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return AERO3D_E_NODEV;
  aero3d_batch *b = new aero3d_batch();
  b->max_slots = max_slots;
  b->slots.resize(max_slots);
  if (!b->arena.init((size_t)arena_bytes) || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void **)&b->d_hdr, sizeof(TableHdr) * max_slots) != hipSuccess ||
      hipMalloc((void **)&b->d_table, sizeof(SlotDev) * (size_t)max_slots * max_slots) != hipSuccess) {
    aero3d_batch_destroy(b);
    return AERO3D_E_HIP_BASE;
  }
  int dev = 0, ncu = 0, occ = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_kernel<1, 1>, 64 * kTY, 0) != hipSuccess) {
    aero3d_batch_destroy(b);
    return AERO3D_E_HIP_BASE;
  }
  b->grid = std::max(1, ncu * std::max(1, occ));       // resident blocks: the persistent form's grid
  *out = b;
  return 0;
}

int aero3d_batch_destroy(aero3d_batch *b) {
  if (!b) return 0;
  // teardown: an error here leaves nothing to recover, so return codes are not checked
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto e : b->execs) if (e) (void)hipGraphExecDestroy(e);
  for (auto g : b->graphs) if (g) (void)hipGraphDestroy(g);
  if (b->arena.base) (void)hipFree(b->arena.base);
  if (b->d_hdr) (void)hipFree(b->d_hdr);
  if (b->d_table) (void)hipFree(b->d_table);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

const char *aero3d_batch_last_error(const aero3d_batch *b) { return b ? b->err.c_str() : ""; }

int aero3d_batch_add_geometry(aero3d_batch *b, const aero3d_grid *g, const uint8_t *t, const float *phi, const float *ap0,
                              const float *ap1, const float *ap2, const float *ap3, const int32_t *fq, const double *fcx,
                              const double *fcz, const double *fx, const double *fz, int nf, int *gid) {
  if (!b || !g || !t || !phi || !ap0 || !ap1 || !ap2 || !ap3 || !gid || nf < 0 || (nf > 0 && (!fq || !fcx || !fcz || !fx || !fz)))
    return bfail(b, AERO3D_E_ARG, "add_geometry: null argument");
  // as aero3d_create: two rows for the y = 0 mirror, 32-bit element offsets
  if (g->NX < 1 || g->NY < 2 || g->NZ < 1 || (int64_t)g->NX * g->NY * g->NZ > (int64_t)1 << 30)
    return bfail(b, AERO3D_E_ARG, "add_geometry: grid size out of range (NX, NZ >= 1, NY >= 2, NX*NY*NZ <= 2^30)");
  Geometry G;
  G.NX = g->NX; G.NY = g->NY; G.NZ = g->NZ; G.N = g->NX * g->NY * g->NZ;
  for (int k = 0; k < nf; k++) if (fq[k] < 0 || fq[k] >= G.N) return bfail(b, AERO3D_E_ARG, "add_geometry: force-list index out of range");
  const size_t N = G.N;
  auto al = [](size_t x) { return (x + kAlign - 1) / kAlign * kAlign; };
  const size_t off_dx = 0, off_dy = off_dx + al(4 * G.NX), off_dz = off_dy + al(4 * G.NY), off_t = off_dz + al(4 * G.NZ),
               off_phi = off_t + al(N), off_ap = off_phi + al(4 * N), off_fq = off_ap + al(16 * N),
               off_fd = off_fq + al(4 * (size_t)nf), total = off_fd + al(32 * (size_t)nf);
  char *m = (char *)b->arena.get(total);
  if (!m) return bfail(b, AERO3D_E_STATE, "add_geometry: arena full");
  G.mem = m; G.bytes = total;
  G.dx = (float *)(m + off_dx); G.dy = (float *)(m + off_dy); G.dz = (float *)(m + off_dz); G.t = (uint8_t *)(m + off_t);
  G.phi = (float *)(m + off_phi); G.ap = (float *)(m + off_ap);
  G.nf = nf; G.fq = (int32_t *)(m + off_fq);
  G.fcx = (double *)(m + off_fd); G.fcz = G.fcx + nf; G.fx = G.fcz + nf; G.fz = G.fx + nf;
  std::vector<float> hdx(G.NX);
  G.hdy.resize(G.NY); G.hdz.resize(G.NZ); G.hxc.resize(G.NX);
  for (int i = 0; i < G.NX; i++) { hdx[i] = (float)(g->fx[i + 1] - g->fx[i]); G.hxc[i] = (float)(0.5 * (g->fx[i] + g->fx[i + 1])); }
  for (int j = 0; j < G.NY; j++) G.hdy[j] = (float)(g->fy[j + 1] - g->fy[j]);
  for (int k = 0; k < G.NZ; k++) G.hdz[k] = (float)(g->fz[k + 1] - g->fz[k]);
  auto up = [&](void *d, const void *h, size_t n) { return hipMemcpy(d, h, n, hipMemcpyHostToDevice); };
  if (up(G.dx, hdx.data(), 4 * G.NX) || up(G.dy, G.hdy.data(), 4 * G.NY) || up(G.dz, G.hdz.data(), 4 * G.NZ) ||
      up(G.t, t, N) || up(G.phi, phi, 4 * N) || up(G.ap, ap0, 4 * N) || up(G.ap + N, ap1, 4 * N) || up(G.ap + 2 * N, ap2, 4 * N) ||
      up(G.ap + 3 * N, ap3, 4 * N) ||
      (nf > 0 && (up(G.fq, fq, 4 * (size_t)nf) || up(G.fcx, fcx, 8 * (size_t)nf) || up(G.fcz, fcz, 8 * (size_t)nf) ||
                  up(G.fx, fx, 8 * (size_t)nf) || up(G.fz, fz, 8 * (size_t)nf)))) {
    b->arena.put(G.mem, G.bytes);
    return bfail(b, AERO3D_E_HIP_BASE, "add_geometry: upload failed");
  }
  for (int k = 0; k < (int)b->geos.size(); k++)
    if (b->geos[k].dead && !b->geos[k].mem) { b->geos[k] = std::move(G); *gid = k; return 0; }
  b->geos.push_back(std::move(G));
  *gid = (int)b->geos.size() - 1;
  return 0;
}

int aero3d_batch_remove_geometry(aero3d_batch *b, int gid) {
  if (!b || gid < 0 || gid >= (int)b->geos.size() || b->geos[gid].dead) return bfail(b, AERO3D_E_ARG, "remove_geometry: bad id");
  Geometry &G = b->geos[gid];
  if (G.refs > 0) return bfail(b, AERO3D_E_STATE, "remove_geometry: still used by a case");
  BCHK(b, hipStreamSynchronize(b->stream));
  b->arena.put(G.mem, G.bytes);
  G.mem = nullptr; G.dead = true;
  return 0;
}

int aero3d_batch_case_bytes(const aero3d_batch *b, int gid, int64_t *bytes) {
  if (!b || !bytes || gid < 0 || gid >= (int)b->geos.size() || b->geos[gid].dead) return AERO3D_E_ARG;
  *bytes = (int64_t)slot_bytes(b->geos[gid], nullptr);
  return 0;
}

int aero3d_batch_add(aero3d_batch *b, int gid, const double F[5], const double fix[15], double cfl, int order, int *slot) {
  if (!b || !F || !fix || !slot || gid < 0 || gid >= (int)b->geos.size() || b->geos[gid].dead) return bfail(b, AERO3D_E_ARG, "add: bad argument");
  if (!(cfl > 0 && cfl < 10) || (order != 1 && order != 2)) return bfail(b, AERO3D_E_ARG, "add: cfl in (0,10), order 1 or 2");
  int s = -1;
  for (int k = 0; k < b->max_slots; k++) if (!b->slots[k].used) { s = k; break; }
  if (s < 0) return bfail(b, AERO3D_E_STATE, "add: no free slot");
  Geometry &G = b->geos[gid];
  size_t ns;
  const size_t bytes = slot_bytes(G, &ns);
  char *m = (char *)b->arena.get(bytes);
  if (!m) return bfail(b, AERO3D_E_STATE, "add: arena full");
  Slot &S = b->slots[s];
  S = Slot();
  S.used = true; S.geo = gid; S.mem = m; S.bytes = bytes; S.ns = ns;
  auto al = [](size_t x) { return (x + kAlign - 1) / kAlign * kAlign; };
  S.state = (float *)m; S.rsnap = (float *)(m + al(4 * 15 * ns)); S.partials = (double *)(m + al(4 * 15 * ns) + al(4 * (size_t)G.N));
  double w[5];
  prim5h(F, w);
  for (int v = 0; v < 5; v++) S.kc.Fw[v] = (float)w[v];
  for (int gg = 0; gg < 3; gg++) {
    prim5h(fix + 5 * gg, w);
    for (int v = 0; v < 5; v++) { S.kc.fixw[gg][v] = (float)w[v]; S.kc.fixq[gg][v] = (float)fix[5 * gg + v]; }
  }
  S.kc.cfl = (float)cfl; S.kc.order = order;
  S.zc = b->zc;
  G.refs++;
  // reset(): state := F
  bfill_kernel<<<(G.N + 255) / 256, 256, 0, b->stream>>>(S.state, ns, G.N, (float)F[0], (float)F[1], (float)F[2], (float)F[3], (float)F[4]);
  BCHK(b, hipGetLastError());
  BFixStates fs;
  for (int gg = 0; gg < 3; gg++)
    for (int v = 0; v < 5; v++) fs.q[gg][v] = S.kc.fixq[gg][v];
  bfixcells_kernel<<<(G.N + 255) / 256, 256, 0, b->stream>>>(S.state, ns, G.N, G.t, fs);
  BCHK(b, hipGetLastError());
  b->table_dirty = true;
  *slot = s;
  return 0;
}

int aero3d_batch_remove(aero3d_batch *b, int slot) {
  if (!b || slot < 0 || slot >= b->max_slots || !b->slots[slot].used) return bfail(b, AERO3D_E_ARG, "remove: bad slot");
  BCHK(b, hipStreamSynchronize(b->stream));
  Slot &S = b->slots[slot];
  b->arena.put(S.mem, S.bytes);
  b->geos[S.geo].refs--;
  S = Slot();
  b->table_dirty = true;
  return 0;
}

int aero3d_batch_set_zchunk(aero3d_batch *b, int zc) {
  if (!b || zc < 1) return AERO3D_E_ARG;
  b->zc = zc;
  for (auto &S : b->slots) if (S.used) S.zc = zc;
  b->table_dirty = true;
  return 0;
}

int aero3d_batch_set_persistent(aero3d_batch *b, int on) {
  if (!b || (on != 0 && on != 1)) return AERO3D_E_ARG;
  b->persistent = on;            // the next step re-captures graphs whose grid changes
  return 0;
}

int aero3d_batch_set_group_cells(aero3d_batch *b, int64_t cells) {
  if (!b || cells < 1) return AERO3D_E_ARG;
  b->group_cells = cells;
  b->table_dirty = true;
  return 0;
}

int aero3d_batch_step(aero3d_batch *b, int n) {
  if (!b || n < 0) return bfail(b, AERO3D_E_ARG, "step: n >= 0");
  return batch_steps(b, n);
}

int aero3d_batch_sync(aero3d_batch *b) {
  if (!b) return AERO3D_E_ARG;
  BCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

int aero3d_batch_iter(const aero3d_batch *b, int slot, int *it) {
  if (!b || !it || slot < 0 || slot >= b->max_slots || !b->slots[slot].used) return AERO3D_E_ARG;
  *it = b->slots[slot].it;
  return 0;
}

int aero3d_batch_read(aero3d_batch *b, int slot, float *r, float *ru, float *rv, float *rw, float *E) {
  if (!b || slot < 0 || slot >= b->max_slots || !b->slots[slot].used) return bfail(b, AERO3D_E_ARG, "read: bad slot");
  Slot &S = b->slots[slot];
  const size_t N = b->geos[S.geo].N;
  float *dst[5] = {r, ru, rv, rw, E};
  for (int c = 0; c < 5; c++)
    if (dst[c]) BCHK(b, hipMemcpyAsync(dst[c], slot_Q(S) + c * S.ns, 4 * N, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return 0;
}

// Residuals of `n` slots (density RMS change per iteration since each slot's
// last snapshot, src/air_cfd.js:553-554); NaN for a slot's first call.
int aero3d_batch_residuals(aero3d_batch *b, const int *slots, int n, double *res) {
  if (!b || n < 0 || (n > 0 && (!slots || !res))) return AERO3D_E_ARG;
  std::vector<int> nit(n);
  for (int k = 0; k < n; k++) {
    if (slots[k] < 0 || slots[k] >= b->max_slots || !b->slots[slots[k]].used) return bfail(b, AERO3D_E_ARG, "residuals: bad slot");
    Slot &S = b->slots[slots[k]];
    nit[k] = (S.snap_it >= 0 && S.it > S.snap_it) ? S.it - S.snap_it : 0;
    enqueue_residual(b, S);
  }
  BCHK(b, hipGetLastError());
  std::vector<double> v(2 * (size_t)n, 0.0);
  for (int k = 0; k < n; k++)
    if (nit[k]) BCHK(b, hipMemcpyAsync(&v[2 * k], b->slots[slots[k]].partials + 3 * kRedMaxBlocks, 16, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  for (int k = 0; k < n; k++) {
    Slot &S = b->slots[slots[k]];
    res[k] = nit[k] ? std::sqrt(v[2 * k] / std::max(1.0, v[2 * k + 1])) / nit[k] : NAN;
    S.snap_it = S.it;
  }
  return 0;
}

int aero3d_batch_forces(aero3d_batch *b, int slot, double alphaDeg, double M, double S_, double xcg, double zcg, double cref, double out[3]) {
  if (!b || !out || slot < 0 || slot >= b->max_slots || !b->slots[slot].used) return bfail(b, AERO3D_E_ARG, "forces: bad slot");
  Slot &S = b->slots[slot];
  enqueue_forces(b, S, xcg, zcg);
  BCHK(b, hipGetLastError());
  double F3[3];
  BCHK(b, hipMemcpyAsync(F3, S.partials + 3 * kRedMaxBlocks + 2, 24, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  coefficients(F3, alphaDeg, M, S_, cref, out);
  return 0;
}

int aero3d_batch_footprint(const aero3d_batch *b, int64_t *used, int64_t *capacity) {
  if (!b || !used || !capacity) return AERO3D_E_ARG;
  *used = (int64_t)b->arena.used; *capacity = (int64_t)b->arena.cap;
  return 0;
}

int aero3d_batch_time_steps(aero3d_batch *b, int n, float *ms) {
  if (!b || !ms || n < 0) return AERO3D_E_ARG;
  int rc;
  if (b->table_dirty && (rc = upload_table(b))) return rc;
  struct Ev { hipEvent_t e = nullptr; ~Ev() { if (e) (void)hipEventDestroy(e); } } e0, e1;   // released on every path
  BCHK(b, hipEventCreate(&e0.e));
  BCHK(b, hipEventCreate(&e1.e));
  BCHK(b, hipEventRecord(e0.e, b->stream));
  if ((rc = batch_steps(b, n))) return rc;
  BCHK(b, hipEventRecord(e1.e, b->stream));
  BCHK(b, hipEventSynchronize(e1.e));
  BCHK(b, hipEventElapsedTime(ms, e0.e, e1.e));
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Server: continuous batching over the batch
// ---------------------------------------------------------------------------
namespace {

struct Pending {
  aero3d_request r;
  double arrival;
};
struct Active {
  aero3d_request r;
  int slot;
  double arrival, t_admit, res0 = NAN, res = NAN;
};

}  // namespace

struct aero3d_server {
  aero3d_batch *b = nullptr;
  int check_every = 50;
  std::deque<Pending> pending;
  std::vector<Active> active;
  std::vector<aero3d_result> done;
  std::chrono::steady_clock::time_point epoch;
  double busy_s = 0;
  int64_t cell_updates = 0;
  std::string err;
};

namespace {

double clock_s(const aero3d_server *s) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->epoch).count();
}

// arena bytes free with no case resident (the registered geometries stay)
int64_t free_when_idle(const aero3d_batch *b) {
  size_t geo = 0;
  for (const auto &G : b->geos)
    if (!G.dead && G.mem) geo += (G.bytes + kAlign - 1) / kAlign * kAlign;
  return (int64_t)b->arena.cap - (int64_t)geo;
}

}  // namespace

extern "C" {

int aero3d_server_create(int64_t arena_bytes, int max_slots, int check_every, aero3d_server **out) {
  if (!out || check_every < 1) return AERO3D_E_ARG;
  *out = nullptr;
  aero3d_server *s = new aero3d_server();
  int rc = aero3d_batch_create(arena_bytes, max_slots, &s->b);
  if (rc) { delete s; return rc; }
  s->check_every = check_every;
  s->epoch = std::chrono::steady_clock::now();
  *out = s;
  return 0;
}

int aero3d_server_destroy(aero3d_server *s) {
  if (!s) return 0;
  aero3d_batch_destroy(s->b);
  delete s;
  return 0;
}

aero3d_batch *aero3d_server_batch(aero3d_server *s) { return s ? s->b : nullptr; }

int aero3d_server_clock(const aero3d_server *s, double *now) {
  if (!s || !now) return AERO3D_E_ARG;
  *now = clock_s(s);
  return 0;
}

int aero3d_server_submit(aero3d_server *s, const aero3d_request *r) {
  if (!s || !r) return AERO3D_E_ARG;
  if (r->geometry < 0 || r->geometry >= (int)s->b->geos.size() || s->b->geos[r->geometry].dead) return AERO3D_E_ARG;
  if (r->max_iters < 1 || !(r->res_drop > 0)) return AERO3D_E_ARG;
  int64_t need;
  int rc = aero3d_batch_case_bytes(s->b, r->geometry, &need);
  if (rc) return rc;
  if (free_when_idle(s->b) < need + (int64_t)kAlign) {
    return bfail(s->b, AERO3D_E_STATE, "submit: request " + std::to_string(r->tag) + " needs " + std::to_string(need) +
                                           " bytes, more than the arena holds beside the registered geometries");
  }
  s->pending.push_back(Pending{*r, r->arrival});
  return 0;
}

// Admit pending requests (FIFO) while the arena has room, step every resident
// case by check_every iterations, measure residuals, retire finished cases.
int aero3d_server_tick(aero3d_server *s, int *n_active, int *n_pending) {
  if (!s) return AERO3D_E_ARG;
  aero3d_batch *b = s->b;
  int rc;
  while (!s->pending.empty()) {
    Pending &p = s->pending.front();
    int64_t need;
    if ((rc = aero3d_batch_case_bytes(b, p.r.geometry, &need))) return rc;
    int slot = -1;
    if ((int64_t)(b->arena.cap - b->arena.used) >= need + (int64_t)kAlign) {      // admission by HBM footprint
      rc = aero3d_batch_add(b, p.r.geometry, p.r.F, p.r.fix, p.r.cfl, p.r.order, &slot);
      if (rc && rc != AERO3D_E_STATE) return rc;
      if (rc) slot = -1;                                                          // no slot / fragmented
    }
    if (slot < 0) {
      if (!s->active.empty()) break;                 // wait: a resident case will free room
      // nothing resident will free room (geometries registered after the
      // submit, or slots taken on the batch directly): return the request as
      // failed so the queue does not stall
      aero3d_result o;
      std::memset(&o, 0, sizeof o);
      o.tag = p.r.tag; o.res0 = o.res = o.CL = o.CD = o.Cm = NAN;
      o.arrival = p.arrival; o.t_admit = o.t_done = clock_s(s);
      s->done.push_back(o);
      b->err = "tick: request " + std::to_string(p.r.tag) + " cannot be admitted with no case resident";
      s->pending.pop_front();
      continue;
    }
    Active a;
    a.r = p.r; a.slot = slot; a.arrival = p.arrival; a.t_admit = clock_s(s);
    s->active.push_back(a);
    s->pending.pop_front();
    // first snapshot (reset state) so the first check measures iterations 0..check_every
    double dummy;
    if ((rc = aero3d_batch_residuals(b, &slot, 1, &dummy))) return rc;
  }
  if (!s->active.empty()) {
    const double t0 = clock_s(s);
    if ((rc = aero3d_batch_step(b, s->check_every))) return rc;
    std::vector<int> sl;
    for (auto &a : s->active) sl.push_back(a.slot);
    std::vector<double> res(sl.size());
    if ((rc = aero3d_batch_residuals(b, sl.data(), (int)sl.size(), res.data()))) return rc;
    for (auto &a : s->active) s->cell_updates += (int64_t)b->geos[a.r.geometry].N * s->check_every;
    std::vector<Active> keep;
    for (size_t k = 0; k < s->active.size(); k++) {
      Active &a = s->active[k];
      a.res = res[k];
      if (!(a.res0 > 0) && a.res > 0) a.res0 = a.res;
      int it;
      aero3d_batch_iter(b, a.slot, &it);
      const bool conv = a.res0 > 0 && a.res / a.res0 <= a.r.res_drop;
      if (conv || it >= a.r.max_iters || !std::isfinite(a.res)) {
        aero3d_result o;
        std::memset(&o, 0, sizeof o);
        o.tag = a.r.tag; o.iters = it; o.res0 = a.res0; o.res = a.res; o.converged = conv ? 1 : 0;
        double c[3] = {NAN, NAN, NAN};
        if (b->geos[a.r.geometry].nf > 0 && (rc = aero3d_batch_forces(b, a.slot, a.r.alpha, a.r.M, a.r.S, a.r.xcg, a.r.zcg, a.r.cref, c))) return rc;
        o.CL = c[0]; o.CD = c[1]; o.Cm = c[2];
        o.arrival = a.arrival; o.t_admit = a.t_admit; o.t_done = clock_s(s);
        if ((rc = aero3d_batch_remove(b, a.slot))) return rc;
        s->done.push_back(o);
      } else {
        keep.push_back(a);
      }
    }
    s->active.swap(keep);
    s->busy_s += clock_s(s) - t0;
  }
  if (n_active) *n_active = (int)s->active.size();
  if (n_pending) *n_pending = (int)s->pending.size();
  return 0;
}

int aero3d_server_poll(aero3d_server *s, aero3d_result *out, int max, int *n) {
  if (!s || !n || max < 0 || (max > 0 && !out)) return AERO3D_E_ARG;
  const int k = std::min(max, (int)s->done.size());
  for (int i = 0; i < k; i++) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return 0;
}

int aero3d_server_stats(const aero3d_server *s, double *busy_s, int64_t *cell_updates) {
  if (!s || !busy_s || !cell_updates) return AERO3D_E_ARG;
  *busy_s = s->busy_s; *cell_updates = s->cell_updates;
  return 0;
}

}  // extern "C"
