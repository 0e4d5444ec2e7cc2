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
//     tile chunks) it owns; it is staged through pinned host buffers and
//     copied asynchronously (no wait on the stream);
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
// as memory frees up (continuous batching). Each tick is enqueued without
// waiting: the steps to the next readback of any case, then the readback of
// every case due (density residual and forces through the shared segment
// reductions) and the convergence rule, decided on the device: a finished case is frozen there
// (its work items exit at once) until the host, which lags one tick behind,
// retires it and frees its slot. The host keeps two ticks in flight and
// waits only for the older one's results.
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
#include "converge.hpp"
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
  int slot;                         // the case's status record
};
struct TableHdr { int nslots, total, base; };   // base: first entry of the group in the table

// Device status of a slot: the served form's convergence state and last
// readback; `frozen` makes the stage kernels skip the case.
struct SlotStat {
  ConvState cs;
  int frozen, it;
  double res, CL, CD, Cm;
};

// Stage kernel over the case table; a work item is (case, tile column, z
// chunk). persist = 0: one item per block, the launch grid is the group's item
// count, so each XCD takes a contiguous range of items. persist = 1: the
// resident blocks walk all items (one persistent launch whatever the table);
// the loop state stays live across the tile stage and spills to scratch, so on
// 16 cases of the batched workload it runs 11.2 against 13.1 G cell-updates/s.
// phase: iterations since the table's buffer indices were written (mod 3);
// stage: 0 or 1 of SSP-RK2. A frozen case's items return at once. At least
// four waves per SIMD (<= 128 VGPRs).
// TX: tile width of every case of the group (64, or 32 for cases whose x
// extent fills 32-column tiles better).
template <int stage, int persist, int TX>
__global__ void __launch_bounds__(TX * kTY) __attribute__((amdgpu_waves_per_eu(4)))
batch_kernel(const TableHdr *__restrict__ hdr, const SlotDev *__restrict__ table, const SlotStat *__restrict__ stat, int phase) {
  __shared__ TileSmem<TX, kTY> sm;
  const int total = hdr->total, nsl = hdr->nslots;
  const SlotDev *__restrict__ slots = table + hdr->base;
  const int G = gridDim.x, r = xcd_remap(blockIdx.x, G);
  auto run = [&](int item) {
    int lo = 0, hi = nsl - 1;
    while (lo < hi) { const int m = (lo + hi + 1) >> 1; if (slots[m].item0 <= item) lo = m; else hi = m - 1; }
    lo = __builtin_amdgcn_readfirstlane(lo);      // block-uniform: keep the case record in scalar registers
    const SlotDev &S = slots[lo];
    if (__builtin_amdgcn_readfirstlane(stat[S.slot].frozen)) return;
    const int li = item - S.item0;
    const int bx = li % S.ntx, by = (li / S.ntx) % S.nty, bz = li / (S.ntx * S.nty);
    const int k0 = bz * S.zc, k1 = min(S.cv.NZ, k0 + S.zc);
    const int q = (S.cur + 2 * phase) % 3;
    const size_t plane5 = 5 * (size_t)S.ns;
    float *Q = S.state + (size_t)q * plane5, *S1 = S.state + (size_t)((q + 1) % 3) * plane5,
          *Bb = S.state + (size_t)((q + 2) % 3) * plane5;
    if constexpr (stage == 0) tile_stage<TX, kTY, false, 1>(S.cv, S.kc, Q, Q, S1, nullptr, S.ns, 0.f, 1.f, bx * TX, by * kTY, k0, k1, sm);
    else tile_stage<TX, kTY, false, 2>(S.cv, S.kc, S1, Q, Bb, nullptr, S.ns, 0.5f, 0.5f, bx * TX, by * kTY, k0, k1, sm);
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
// A new case's status record: no readback yet, not frozen.
__global__ void stat_reset_kernel(SlotStat *__restrict__ stat, int slot) {
  if (threadIdx.x != 0) return;
  SlotStat &x = stat[slot];
  conv_init(x.cs);
  x.frozen = 0; x.it = 0;
  x.res = x.CL = x.CD = x.Cm = NAN;
}
// Served form: final sums of each due case, then its convergence rule; a case
// that is done is frozen. A case frozen earlier keeps its record.
__global__ void __launch_bounds__(kRedBlock) srv_finish_kernel(const RbSeg *__restrict__ seg, SlotStat *__restrict__ stat) {
  __shared__ double l2[kRedBlock / 64][2];
  __shared__ double l3[kRedBlock / 64][3];
  const RbSeg g = seg[blockIdx.x];
  SlotStat &x = stat[g.slot];
  if (__builtin_amdgcn_readfirstlane(x.frozen)) return;
  finish_sum<2>(g.part, g.nb_res, g.out, l2);
  finish_sum<3>(g.part + 2 * g.nb_res, g.nb_f, g.out + 2, l3);
  if (threadIdx.x == 0) {
    rb_results(g);
    const double *o = g.out;
    conv_update(x.cs, g.rule, g.res_drop, g.cap, o[5], o[6], g.it);
    x.it = g.it; x.res = o[5]; x.CL = o[6]; x.CD = o[7]; x.Cm = o[8];
    if (x.cs.done) x.frozen = 1;
  }
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
  double *partials = nullptr;       // FP64 readback partials (5 kRedMaxBlocks) and results (kRbOut)
  size_t ns = 0;
  int cur = 0, it = 0, snap_it = -1;
  StageConsts kc;
  int zc = 3;
};

// Pinned host staging buffers, reused in turn; each records an event after
// the last copy that reads or writes it, and is waited for before reuse.
struct PinnedRing {
  static constexpr int kN = 6;
  char *buf[kN] = {};
  hipEvent_t ev[kN] = {};
  bool held[kN] = {};
  size_t cap = 0;
  int next = 0;
};

}  // namespace

struct aero3d_batch {
  Arena arena;
  hipStream_t stream = nullptr;
  std::vector<Geometry> geos;
  std::vector<Slot> slots;
  TableHdr *d_hdr = nullptr;         // one header per case group
  SlotDev *d_table = nullptr;        // max_slots entries, groups one after another
  SlotStat *d_stat = nullptr;        // max_slots status records
  RbSeg *d_seg = nullptr;            // 2 x max_slots readback segments (admission, check)
  int max_slots = 0;
  int grid = 0, grid32 = 0;          // resident blocks of the persistent form (64- / 32-column tiles)
  std::vector<hipGraph_t> graphs;    // per case group: three iterations
  std::vector<hipGraphExec_t> execs;
  std::vector<int> totals;           // per case group: work items (the launch grid)
  std::vector<int> gtx;              // per case group: tile width of its cases
  int tile_width = 0;                // 64 or 32 for every case; 0: per case (case_tx)
  std::vector<int> graph_total;      // per case group: the launch grid and tile width its graph was captured with
  int persistent = 0;                // 1: resident blocks walk the items (batch_kernel<., 1>)
  bool table_dirty = true;
  std::vector<int> order;            // slot ids in table order
  int ngroups = 0;
  int64_t group_cells = int64_t(1) << 40;  // cases are stepped in groups of about this many cells (default: one group)
  int zc = 0;                        // z chunk of every case; 0: from the load (load_zc)
  PinnedRing ring;
  std::string err;
};

namespace {

int bfail(aero3d_batch *b, int code, const std::string &m) { if (b) b->err = m; return code; }
#define BCHK(b, expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return bfail((b), AERO3D_E_HIP_BASE - (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Take the next staging buffer (waiting for its previous copies if they are
// still pending); *k = its index. The caller records ring.ev[*k] after its
// last copy and releases it.
int ring_acquire(aero3d_batch *b, int *k) {
  PinnedRing &R = b->ring;
  for (int n = 0; n < PinnedRing::kN; n++) {
    const int i = (R.next + n) % PinnedRing::kN;
    if (R.held[i]) continue;
    BCHK(b, hipEventSynchronize(R.ev[i]));
    R.held[i] = true;
    R.next = (i + 1) % PinnedRing::kN;
    *k = i;
    return 0;
  }
  return bfail(b, AERO3D_E_STATE, "staging buffers exhausted");
}
int ring_release(aero3d_batch *b, int k) {
  BCHK(b, hipEventRecord(b->ring.ev[k], b->stream));
  b->ring.held[k] = false;
  return 0;
}

size_t slot_bytes(const Geometry &g, size_t *ns_out) {
  const size_t ns = ((size_t)g.N + 63) / 64 * 64;
  if (ns_out) *ns_out = ns;
  auto al = [](size_t b) { return (b + kAlign - 1) / kAlign * kAlign; };
  return al(4 * 15 * ns) + al(4 * (size_t)g.N) + al(8 * (5 * kRedMaxBlocks + 16));
}

void prim5h(const double *a, double *o) {
  const double r = std::max(a[0], 1e-6), u = a[1] / r, v = a[2] / r, w = a[3] / r;
  o[0] = r; o[1] = u; o[2] = v; o[3] = w;
  o[4] = std::max((1.4 - 1) * (a[4] - 0.5 * r * (u * u + v * v + w * w)), 1e-7);
}

// Byte offsets inside one staging buffer: table headers, table rows, two
// segment areas, the slot status copy.
struct StageLayout {
  size_t hdr, rows, segA, segC, stat, total;
  explicit StageLayout(int m) {
    auto al = [](size_t x) { return (x + kAlign - 1) / kAlign * kAlign; };
    hdr = 0; rows = al(sizeof(TableHdr) * m); segA = rows + al(sizeof(SlotDev) * m); segC = segA + al(sizeof(RbSeg) * m);
    stat = segC + al(sizeof(RbSeg) * m); total = stat + al(sizeof(SlotStat) * m);
  }
};

// Rewrite the device case table from the active slots, through a pinned
// staging buffer (asynchronous). Cases are split into groups of about
// group_cells cells (a group's state and masks stay in the Infinity Cache
// across its stages and iterations); each group has its own header and a
// contiguous range of the table, stepped by its own launches.
// z chunk of every case when the batch's is 0: the longest of 8, 6, 4 and 3
// planes that still makes at least four work items per resident block. Long
// chunks pay the z queue's preamble (the planes below and above a chunk) over
// more planes, short ones give the cases in flight more items. Measured on the
// served request mix: with all 200 requests at once 8 planes give 53.5
// solves/s against 48.5 at 3; the arrival-bound stream (a few requests in
// flight) runs at p50 / p99 0.13 / 0.37 s with 3 planes against 0.20 / 0.41 s
// with 8.
// Tile width of a case: 32 when 32-column tiles cover its x extent with fewer
// padded columns than 64-column ones (e.g. NX 72, 95, 160), else 64; or the
// batch's fixed width. The state is the same bit for bit either way. Measured
// on the plain path (one case, us per iteration, 64 / 32 wide): NX 72 22.1 /
// 21.5, 95 22.2 / 21.0, 160 42.2 / 36.5; where the padded widths are equal,
// 64 wins (NX 115: 40.7 / 43.2, NX 174: 60.8 / 62.9).
int case_tx(const aero3d_batch *b, int NX) {
  if (b->tile_width) return b->tile_width;
  return (NX + 31) / 32 * 32 < (NX + 63) / 64 * 64 ? 32 : 64;
}

int load_zc(const aero3d_batch *b) {
  const int cand[4] = {8, 6, 4, 3};
  for (int c = 0; c < 4; c++) {
    int64_t items = 0;
    for (const Slot &S : b->slots) {
      if (!S.used) continue;
      const Geometry &g = b->geos[S.geo];
      const int tx = case_tx(b, g.NX);
      items += (int64_t)((g.NX + tx - 1) / tx) * ((g.NY + kTY - 1) / kTY) * ((g.NZ + cand[c] - 1) / cand[c]);
    }
    if (items >= 4 * (int64_t)b->grid) return cand[c];
  }
  return 3;
}

// Per-case widths (tile_width 0) split the cases into one launch per width per
// stage. That pays only when each launch still fills the device: with fewer
// than two work items per resident block in either width, every case takes
// 64-column tiles (one launch). Measured on the served stream (a few requests in
// flight): split, p50 0.197 s and p99 0.399 s; one launch, 0.130 s and 0.370 s.
bool split_widths(const aero3d_batch *b, int zc) {
  if (b->tile_width) return true;
  int64_t it[2] = {0, 0};
  for (const Slot &S : b->slots) {
    if (!S.used) continue;
    const Geometry &g = b->geos[S.geo];
    const int tx = case_tx(b, g.NX);
    it[tx == 32] += (int64_t)((g.NX + tx - 1) / tx) * ((g.NY + kTY - 1) / kTY) * ((g.NZ + zc - 1) / zc);
  }
  return it[0] == 0 || it[1] == 0 || (it[0] >= 2 * (int64_t)b->grid && it[1] >= 2 * (int64_t)b->grid32);
}

int upload_table(aero3d_batch *b) {
  b->order.clear();
  const int zc_load = b->zc > 0 ? b->zc : load_zc(b);
  const bool split = split_widths(b, zc_load);
  std::vector<SlotDev> rows;
  std::vector<TableHdr> hdrs;
  int64_t cells = 0;
  b->gtx.clear();
  // cases of 64-column tiles first, then those of 32-column tiles: a group holds one tile width
  for (int pass = 0; pass < 2; pass++)
  for (int s = 0; s < (int)b->slots.size(); s++) {
    Slot &S = b->slots[s];
    if (!S.used) continue;
    const Geometry &g = b->geos[S.geo];
    const int tx = split ? case_tx(b, g.NX) : 64;
    if ((tx == 32) != (pass == 1)) continue;
    if (hdrs.empty() || b->gtx.back() != tx || (hdrs.back().nslots > 0 && cells + g.N > b->group_cells)) {
      hdrs.push_back(TableHdr{0, 0, (int)rows.size()});
      b->gtx.push_back(tx);
      cells = 0;
    }
    cells += g.N;
    SlotDev d;
    d.cv = CaseView{g.NX, g.NY, g.NZ, g.N, g.dx, g.dy, g.dz, g.t, g.phi, g.ap};
    d.kc = S.kc;
    d.state = S.state; d.ns = S.ns; d.cur = S.cur;
    S.zc = zc_load;
    d.zc = S.zc;
    d.ntx = (g.NX + tx - 1) / tx; d.nty = (g.NY + kTY - 1) / kTY;
    d.nitems = d.ntx * d.nty * ((g.NZ + S.zc - 1) / S.zc);
    d.item0 = hdrs.back().total;
    d.slot = s;
    hdrs.back().total += d.nitems;
    hdrs.back().nslots++;
    rows.push_back(d);
    b->order.push_back(s);
  }
  b->ngroups = (int)hdrs.size();
  b->totals.clear();
  for (auto &h : hdrs) b->totals.push_back(h.total);
  if (!rows.empty()) {
    int k, rc;
    if ((rc = ring_acquire(b, &k))) return rc;
    char *p = b->ring.buf[k];
    std::memcpy(p, hdrs.data(), sizeof(TableHdr) * hdrs.size());
    const StageLayout L(b->max_slots);
    std::memcpy(p + L.rows, rows.data(), sizeof(SlotDev) * rows.size());
    BCHK(b, hipMemcpyAsync(b->d_hdr, p, sizeof(TableHdr) * hdrs.size(), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipMemcpyAsync(b->d_table, p + L.rows, sizeof(SlotDev) * rows.size(), hipMemcpyHostToDevice, b->stream));
    if ((rc = ring_release(b, k))) return rc;
  }
  b->table_dirty = false;
  return 0;
}

// Both stages of one SSP-RK2 iteration of case group gi at rotation phase ph.
template <int TX>
void launch_iteration_tx(aero3d_batch *b, int gi, int ph) {
  const TableHdr *h = b->d_hdr + gi;
  if (b->persistent) {
    const int G = TX == 32 ? b->grid32 : b->grid;
    batch_kernel<0, 1, TX><<<G, dim3(TX, kTY), 0, b->stream>>>(h, b->d_table, b->d_stat, ph);
    batch_kernel<1, 1, TX><<<G, dim3(TX, kTY), 0, b->stream>>>(h, b->d_table, b->d_stat, ph);
  } else {
    const int G = std::max(1, b->totals[gi]);
    batch_kernel<0, 0, TX><<<G, dim3(TX, kTY), 0, b->stream>>>(h, b->d_table, b->d_stat, ph);
    batch_kernel<1, 0, TX><<<G, dim3(TX, kTY), 0, b->stream>>>(h, b->d_table, b->d_stat, ph);
  }
}
void launch_iteration(aero3d_batch *b, int gi, int ph) {
  if (b->gtx[gi] == 32) launch_iteration_tx<32>(b, gi, ph);
  else launch_iteration_tx<64>(b, gi, ph);
}

// the launch grid and tile width a group's iteration uses (what its captured graph holds)
int group_grid(const aero3d_batch *b, int gi) {
  const int g = b->persistent ? -(b->gtx[gi] == 32 ? b->grid32 : b->grid) : std::max(1, b->totals[gi]);
  return 2 * g + (b->gtx[gi] == 32 ? 1 : 0);
}

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

// The readback segment of slot s (residual when res, forces when frc). The
// coefficient constants come from (alpha, M, S, cref) as in cfdForces.
RbSeg make_seg(const aero3d_batch *b, int s, bool res, bool frc, double alphaDeg, double M, double Sref, double xcg, double zcg,
               double cref) {
  const Slot &S = b->slots[s];
  const Geometry &g = b->geos[S.geo];
  RbSeg x{};
  x.Q = slot_Q(S); x.ns = S.ns; x.t = g.t; x.N = g.N; x.rsnap = S.rsnap;
  x.nit = (S.snap_it >= 0 && S.it > S.snap_it) ? S.it - S.snap_it : 0;
  x.fq = g.fq; x.cx = g.fcx; x.cz = g.fcz; x.x = g.fx; x.z = g.fz; x.nf = frc ? g.nf : 0;
  x.xcg = xcg; x.zcg = zcg;
  const double a = alphaDeg * (M_PI / 180);
  x.ca = std::cos(a); x.sa = std::sin(a); x.kq = 2 / (0.5 * M * M * Sref); x.cref = cref;
  x.nb_res = res ? reduce_blocks((g.N + 2) / 3) : 0;
  x.nb_f = frc ? reduce_blocks(g.nf) : 0;
  x.slot = s; x.it = S.it;
  x.part = S.partials; x.out = S.partials + 5 * kRedMaxBlocks;
  return x;
}

// Stage segments in staging buffer area `p`, copy them to device segments
// dseg and enqueue their partial sums (the caller enqueues the finish).
int enqueue_segments(aero3d_batch *b, std::vector<RbSeg> &segs, char *p, RbSeg *dseg) {
  int blk = 0;
  for (auto &x : segs) { x.blk0 = blk; blk += x.nb_res + x.nb_f; }
  std::memcpy(p, segs.data(), sizeof(RbSeg) * segs.size());
  BCHK(b, hipMemcpyAsync(dseg, p, sizeof(RbSeg) * segs.size(), hipMemcpyHostToDevice, b->stream));
  if (blk > 0) rb_partials_kernel<<<blk, kRedBlock, 0, b->stream>>>(dseg, (int)segs.size());
  BCHK(b, hipGetLastError());
  return 0;
}

void free_slot(aero3d_batch *b, int slot) {
  Slot &S = b->slots[slot];
  b->arena.put(S.mem, S.bytes);
  b->geos[S.geo].refs--;
  S = Slot();
  b->table_dirty = true;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI: batch
// ---------------------------------------------------------------------------
extern "C" {

int aero3d_batch_create(int64_t arena_bytes, int max_slots, aero3d_batch **out) {
  if (!out || arena_bytes <= 0 || max_slots < 1 || max_slots > 4096) return AERO3D_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return AERO3D_E_NODEV;
  aero3d_batch *b = new aero3d_batch();
  b->max_slots = max_slots;
  b->slots.resize(max_slots);
  const StageLayout L(max_slots);
  bool ok = b->arena.init((size_t)arena_bytes) && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc((void **)&b->d_hdr, sizeof(TableHdr) * max_slots) == hipSuccess &&
            hipMalloc((void **)&b->d_table, sizeof(SlotDev) * (size_t)max_slots) == hipSuccess &&
            hipMalloc((void **)&b->d_stat, sizeof(SlotStat) * (size_t)max_slots) == hipSuccess &&
            hipMalloc((void **)&b->d_seg, sizeof(RbSeg) * 2 * (size_t)max_slots) == hipSuccess;
  b->ring.cap = L.total;
  for (int k = 0; ok && k < PinnedRing::kN; k++)
    ok = hipHostMalloc((void **)&b->ring.buf[k], L.total, hipHostMallocDefault) == hipSuccess &&
         hipEventCreate(&b->ring.ev[k]) == hipSuccess;
  if (!ok) {
    aero3d_batch_destroy(b);
    return AERO3D_E_HIP_BASE;
  }
  int dev = 0, ncu = 0, occ = 0, occ32 = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, batch_kernel<1, 1, 64>, 64 * kTY, 0) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ32, batch_kernel<1, 1, 32>, 32 * kTY, 0) != hipSuccess) {
    aero3d_batch_destroy(b);
    return AERO3D_E_HIP_BASE;
  }
  b->grid = std::max(1, ncu * std::max(1, occ));       // resident blocks: the persistent form's grid
  b->grid32 = std::max(1, ncu * std::max(1, occ32));
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
  void *dp[] = {b->d_hdr, b->d_table, b->d_stat, b->d_seg};
  for (void *p : dp) if (p) (void)hipFree(p);
  for (int k = 0; k < PinnedRing::kN; k++) {
    if (b->ring.buf[k]) (void)hipHostFree(b->ring.buf[k]);
    if (b->ring.ev[k]) (void)hipEventDestroy(b->ring.ev[k]);
  }
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
  // synchronous copies: the caller's arrays may change after return
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
  S.zc = b->zc;                      // 0: chosen from the load when the table is built
  G.refs++;
  // reset(): state := F; fixed cells := their states; status record cleared.
  // Stream-ordered after every earlier launch, so memory of a retired case can be reused at once.
  bfill_kernel<<<(G.N + 255) / 256, 256, 0, b->stream>>>(S.state, ns, G.N, (float)F[0], (float)F[1], (float)F[2], (float)F[3], (float)F[4]);
  BFixStates fs;
  for (int gg = 0; gg < 3; gg++)
    for (int v = 0; v < 5; v++) fs.q[gg][v] = S.kc.fixq[gg][v];
  bfixcells_kernel<<<(G.N + 255) / 256, 256, 0, b->stream>>>(S.state, ns, G.N, G.t, fs);
  stat_reset_kernel<<<1, 64, 0, b->stream>>>(b->d_stat, s);
  BCHK(b, hipGetLastError());
  b->table_dirty = true;
  *slot = s;
  return 0;
}

int aero3d_batch_remove(aero3d_batch *b, int slot) {
  if (!b || slot < 0 || slot >= b->max_slots || !b->slots[slot].used) return bfail(b, AERO3D_E_ARG, "remove: bad slot");
  free_slot(b, slot);
  return 0;
}

int aero3d_batch_set_tile_width(aero3d_batch *b, int width) {
  if (!b || !(width == 0 || width == 32 || width == 64)) return bfail(b, AERO3D_E_ARG, "set_tile_width: 0, 32 or 64");
  b->tile_width = width;
  b->table_dirty = true;
  return 0;
}

int aero3d_batch_set_zchunk(aero3d_batch *b, int zc) {
  if (!b || zc < 0) return AERO3D_E_ARG;
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
  if (!b || n < 0 || n > b->max_slots || (n > 0 && (!slots || !res))) return AERO3D_E_ARG;
  std::vector<RbSeg> segs;
  for (int k = 0; k < n; k++) {
    if (slots[k] < 0 || slots[k] >= b->max_slots || !b->slots[slots[k]].used) return bfail(b, AERO3D_E_ARG, "residuals: bad slot");
    for (int j = 0; j < k; j++) if (slots[j] == slots[k]) return bfail(b, AERO3D_E_ARG, "residuals: slot listed twice");
    segs.push_back(make_seg(b, slots[k], true, false, 0, 1, 1, 0, 0, 1));
  }
  if (n == 0) return 0;
  int k, rc;
  if ((rc = ring_acquire(b, &k))) return rc;
  const StageLayout L(b->max_slots);
  char *p = b->ring.buf[k];
  if ((rc = enqueue_segments(b, segs, p + L.segC, b->d_seg + b->max_slots))) return rc;
  rb_finish_kernel<<<n, kRedBlock, 0, b->stream>>>(b->d_seg + b->max_slots);
  BCHK(b, hipGetLastError());
  std::vector<double> v(n);
  for (int j = 0; j < n; j++) BCHK(b, hipMemcpyAsync(&v[j], segs[j].out + 5, 8, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  if ((rc = ring_release(b, k))) return rc;
  for (int j = 0; j < n; j++) { res[j] = v[j]; b->slots[slots[j]].snap_it = b->slots[slots[j]].it; }
  return 0;
}

int aero3d_batch_forces(aero3d_batch *b, int slot, double alphaDeg, double M, double S_, double xcg, double zcg, double cref, double out[3]) {
  if (!b || !out || slot < 0 || slot >= b->max_slots || !b->slots[slot].used) return bfail(b, AERO3D_E_ARG, "forces: bad slot");
  std::vector<RbSeg> segs{make_seg(b, slot, false, true, alphaDeg, M, S_, xcg, zcg, cref)};
  int k, rc;
  if ((rc = ring_acquire(b, &k))) return rc;
  const StageLayout L(b->max_slots);
  if ((rc = enqueue_segments(b, segs, b->ring.buf[k] + L.segC, b->d_seg + b->max_slots))) return rc;
  rb_finish_kernel<<<1, kRedBlock, 0, b->stream>>>(b->d_seg + b->max_slots);
  BCHK(b, hipGetLastError());
  BCHK(b, hipMemcpyAsync(out, segs[0].out + 6, 24, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return ring_release(b, k);
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

struct SrvCase {                    // a resident request (host view)
  bool live = false;
  unsigned gen = 0;                 // admissions into this slot so far
  aero3d_request r;
  double arrival = 0, t_admit = 0;
  int last_read_it = 0;             // iterations at its last readback
};
struct Due { int slot; unsigned gen; int cur, it; };
struct TickRec {                    // one enqueued tick
  int buf;                          // staging buffer (segments, status copy)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<Due> due;
};

// A request's next readback: every check_every iterations from its admission,
// and, under the residual-drop rule, at its iteration cap.
int next_check(const aero3d_request &r, int it, int every) {
  const int n = (it / every + 1) * every;
  return r.rule == kRuleResDrop ? std::min(n, r.max_iters) : n;
}

}  // namespace

struct aero3d_server {
  aero3d_batch *b = nullptr;
  int check_every = 50;
  std::deque<aero3d_request> pending;
  std::vector<SrvCase> cases;       // by slot
  int live = 0;
  std::vector<aero3d_result> done;
  std::deque<TickRec> inflight;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> evpool;
  std::chrono::steady_clock::time_point epoch;
  double busy_s = 0, gpu_s = 0;
  int64_t cell_updates = 0;
  double t_busy0 = -1;              // start of the current stretch with resident work
};

namespace {

constexpr int kInFlight = 2;        // ticks enqueued ahead of the host's processing

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

aero3d_result failed_result(const aero3d_server *s, const aero3d_request &r, double t) {
  aero3d_result o;
  std::memset(&o, 0, sizeof o);
  o.tag = r.tag; o.failed = 1;
  o.res0 = o.res = o.CL = o.CD = o.Cm = NAN;
  o.arrival = r.arrival; o.t_admit = o.t_done = t;
  (void)s;
  return o;
}

// Results of the oldest enqueued tick: retire every due case the device froze.
int process_oldest(aero3d_server *s) {
  aero3d_batch *b = s->b;
  TickRec T = s->inflight.front();
  s->inflight.pop_front();
  BCHK(b, hipEventSynchronize(T.ev1));
  float ms = 0;
  BCHK(b, hipEventElapsedTime(&ms, T.ev0, T.ev1));
  s->gpu_s += ms * 1e-3;
  s->evpool.push_back({T.ev0, T.ev1});
  const StageLayout L(b->max_slots);
  const SlotStat *st = (const SlotStat *)(b->ring.buf[T.buf] + L.stat);
  const double now = clock_s(s);
  for (const Due &d : T.due) {
    SrvCase &c = s->cases[d.slot];
    if (!c.live || c.gen != d.gen) continue;     // retired at an earlier tick
    const SlotStat &x = st[d.slot];
    if (!x.frozen) continue;
    aero3d_result o;
    std::memset(&o, 0, sizeof o);
    o.tag = c.r.tag; o.iters = x.it; o.converged = x.cs.conv; o.diverged = x.cs.diverged; o.nreads = x.cs.nreads;
    o.res0 = x.cs.res0; o.res = x.res; o.CL = x.CL; o.CD = x.CD; o.Cm = x.Cm;
    o.arrival = c.arrival; o.t_admit = c.t_admit; o.t_done = now;
    if (c.r.state_out) {            // the frozen state, still in buffer d.cur (stream-ordered copy, then a wait)
      const Slot &S = b->slots[d.slot];
      const size_t N = b->geos[S.geo].N;
      const float *Q = S.state + (size_t)d.cur * 5 * S.ns;
      for (int v = 0; v < 5; v++)
        BCHK(b, hipMemcpyAsync(c.r.state_out + v * N, Q + v * S.ns, 4 * N, hipMemcpyDeviceToHost, b->stream));
      BCHK(b, hipStreamSynchronize(b->stream));
    }
    s->cell_updates += (int64_t)x.it * b->geos[b->slots[d.slot].geo].N;
    free_slot(b, d.slot);
    c.live = false;
    s->live--;
    s->done.push_back(o);
  }
  b->ring.held[T.buf] = false;       // its event (T.ev1's copy) is already complete
  if (s->live == 0 && s->t_busy0 >= 0) { s->busy_s += now - s->t_busy0; s->t_busy0 = -1; }
  return 0;
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
  s->cases.resize(max_slots);
  s->epoch = std::chrono::steady_clock::now();
  *out = s;
  return 0;
}

int aero3d_server_destroy(aero3d_server *s) {
  if (!s) return 0;
  if (s->b && s->b->stream) (void)hipStreamSynchronize(s->b->stream);
  for (auto &T : s->inflight) s->evpool.push_back({T.ev0, T.ev1});
  for (auto &e : s->evpool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
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
  aero3d_batch *b = s->b;
  if (r->geometry < 0 || r->geometry >= (int)b->geos.size() || b->geos[r->geometry].dead)
    return bfail(b, AERO3D_E_ARG, "submit: unknown geometry");
  if (!(r->cfl > 0 && r->cfl < 10) || (r->order != 1 && r->order != 2)) return bfail(b, AERO3D_E_ARG, "submit: cfl in (0,10), order 1 or 2");
  if (r->rule != AERO3D_RULE_RES_DROP && r->rule != AERO3D_RULE_REFERENCE) return bfail(b, AERO3D_E_ARG, "submit: unknown rule");
  if (r->max_iters < 1) return bfail(b, AERO3D_E_ARG, "submit: max_iters >= 1");
  if (r->rule == AERO3D_RULE_RES_DROP && !(r->res_drop > 0)) return bfail(b, AERO3D_E_ARG, "submit: res_drop > 0");
  if (!(r->M > 0) || !(r->S > 0) || !(r->cref != 0) || !std::isfinite(r->alpha))
    return bfail(b, AERO3D_E_ARG, "submit: coefficient reference needs M > 0, S > 0, cref != 0, finite alpha");
  int64_t need;
  int rc = aero3d_batch_case_bytes(b, r->geometry, &need);
  if (rc) return rc;
  if (free_when_idle(b) < need + (int64_t)kAlign) {
    return bfail(b, AERO3D_E_STATE, "submit: request " + std::to_string(r->tag) + " needs " + std::to_string(need) +
                                        " bytes, more than the arena holds beside the registered geometries");
  }
  s->pending.push_back(*r);
  return 0;
}

// Process finished ticks (waiting only when kInFlight ticks are queued, or
// for the last ones when nothing is resident), admit pending requests by
// footprint, and enqueue the next tick.
int aero3d_server_tick(aero3d_server *s, int *n_active, int *n_pending) {
  if (!s) return AERO3D_E_ARG;
  aero3d_batch *b = s->b;
  int rc;
  while (!s->inflight.empty()) {
    const bool full = (int)s->inflight.size() >= kInFlight;
    const hipError_t q = hipEventQuery(s->inflight.front().ev1);
    if (q != hipSuccess && q != hipErrorNotReady) BCHK(b, q);
    if (!(full || q == hipSuccess || s->live == 0)) break;
    if ((rc = process_oldest(s))) return rc;
  }
  // admissions (FIFO, by HBM footprint)
  while (!s->pending.empty()) {
    const aero3d_request r = s->pending.front();
    int64_t need = 0;
    int slot = -1;
    rc = aero3d_batch_case_bytes(b, r.geometry, &need);
    if (!rc && (int64_t)(b->arena.cap - b->arena.used) >= need + (int64_t)kAlign) {
      rc = aero3d_batch_add(b, r.geometry, r.F, r.fix, r.cfl, r.order, &slot);
      if (rc == AERO3D_E_STATE) { rc = 0; slot = -1; }                    // no slot / fragmented: wait
    }
    if (rc == AERO3D_E_ARG) {        // cannot run (geometry removed since the submit): a failed result
      s->done.push_back(failed_result(s, r, clock_s(s)));
      b->err = "tick: request " + std::to_string(r.tag) + " rejected at admission";
      s->pending.pop_front();
      continue;
    }
    if (rc) return rc;
    if (slot < 0) {
      if (s->live > 0 || !s->inflight.empty()) break;   // a resident case will free room
      // nothing resident will free room: return the request as failed so the queue does not stall
      s->done.push_back(failed_result(s, r, clock_s(s)));
      b->err = "tick: request " + std::to_string(r.tag) + " cannot be admitted with no case resident";
      s->pending.pop_front();
      continue;
    }
    SrvCase &c = s->cases[slot];
    c.live = true; c.gen++; c.r = r; c.arrival = r.arrival; c.t_admit = clock_s(s); c.last_read_it = 0;
    s->live++;
    s->pending.pop_front();
  }
  if (s->live > 0) {
    if (s->t_busy0 < 0) s->t_busy0 = clock_s(s);
    TickRec T;
    if ((rc = ring_acquire(b, &T.buf))) return rc;
    if (s->evpool.empty()) {
      hipEvent_t e0, e1;
      BCHK(b, hipEventCreate(&e0));
      BCHK(b, hipEventCreate(&e1));
      s->evpool.push_back({e0, e1});
    }
    T.ev0 = s->evpool.back().first; T.ev1 = s->evpool.back().second;
    s->evpool.pop_back();
    const StageLayout L(b->max_slots);
    char *p = b->ring.buf[T.buf];
    BCHK(b, hipEventRecord(T.ev0, b->stream));
    // a case's first readback (no snapshot yet) only takes the snapshot: its
    // residual is NaN and res0 comes from the second, as in the reference's controller
    auto seg_of = [&](int slot) {
      const aero3d_request &r = s->cases[slot].r;
      RbSeg x = make_seg(b, slot, true, true, r.alpha, std::max(r.M, 0.05), r.S, r.xcg, r.zcg, r.cref);
      x.rule = r.rule; x.cap = r.max_iters; x.res_drop = r.res_drop;
      x.skip = &b->d_stat[slot].frozen;
      return x;
    };
    // steps to the next readback of any resident case
    int n = 1 << 30;
    for (int slot = 0; slot < b->max_slots; slot++) {
      const SrvCase &c = s->cases[slot];
      if (c.live) n = std::min(n, next_check(c.r, b->slots[slot].it, s->check_every) - b->slots[slot].it);
    }
    n = std::max(n, 1);
    if ((rc = batch_steps(b, n))) return rc;
    // readback and rule of every case due
    std::vector<RbSeg> segs;
    for (int slot = 0; slot < b->max_slots; slot++) {
      SrvCase &c = s->cases[slot];
      if (!c.live) continue;
      Slot &S = b->slots[slot];
      if (S.it != next_check(c.r, S.it - n, s->check_every)) continue;
      segs.push_back(seg_of(slot));
      S.snap_it = S.it;
      c.last_read_it = S.it;
      T.due.push_back(Due{slot, c.gen, S.cur, S.it});
    }
    if (!segs.empty()) {
      if ((rc = enqueue_segments(b, segs, p + L.segC, b->d_seg + b->max_slots))) return rc;
      srv_finish_kernel<<<(int)segs.size(), kRedBlock, 0, b->stream>>>(b->d_seg + b->max_slots, b->d_stat);
    }
    BCHK(b, hipGetLastError());
    BCHK(b, hipMemcpyAsync(p + L.stat, b->d_stat, sizeof(SlotStat) * b->max_slots, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipEventRecord(T.ev1, b->stream));
    BCHK(b, hipEventRecord(b->ring.ev[T.buf], b->stream));
    s->inflight.push_back(T);          // the buffer stays held until the tick is processed
  }
  if (n_active) *n_active = s->live;
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

int aero3d_server_stats(const aero3d_server *s, double *busy_s, double *gpu_s, int64_t *cell_updates) {
  if (!s || !busy_s || !gpu_s || !cell_updates) return AERO3D_E_ARG;
  *busy_s = s->busy_s + (s->t_busy0 >= 0 ? clock_s(s) - s->t_busy0 : 0);
  *gpu_s = s->gpu_s; *cell_updates = s->cell_updates;
  return 0;
}

}  // extern "C"
