// CPU run of the axisymmetric face / cell kernels (3d_aero_lab_amd/csrc/axi_kernels.hpp)
// for indexing and arithmetic checks under AddressSanitizer. Test infrastructure only.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "../../3d_aero_lab_amd/csrc/axi_kernels.hpp"

thread_local emul_dim3 threadIdx;
thread_local emul_dim3 blockIdx;
thread_local std::barrier<> *emul_block_barrier = nullptr;

using namespace axik;

template <class K>
static void launch(int n, K k) {
  for (int b = 0; b < (n + kB - 1) / kB; b++)
    for (int t = 0; t < kB; t++) { blockIdx.x = b; threadIdx.x = t; k(); }
}

// One launch of the fused tile kernel: every block as 64 NW OS threads sharing
// one AxiTile (NaN-filled, so a read of an unwritten slot shows up) and a barrier.
template <int kMode, int TH, int NW>
static void fused_launch(const AxiView &v, const AxiConsts &c, const float *Qin, const float *Qb, float *Qout, float *dt, float *src,
                         float *mfx) {
  const int ntx = (v.W + kTW - 1) / kTW, nty = (v.H + TH - 1) / TH;
  for (int b = 0; b < ntx * nty; b++) {
    std::barrier<> bar(64 * NW);
    auto tile = std::make_unique<AxiTile<TH>>();
    std::memset(tile.get(), 0xff, sizeof(AxiTile<TH>));
    std::vector<std::thread> th;
    for (int t = 0; t < 64 * NW; t++)
      th.emplace_back([&, t, b] {
        threadIdx.x = t; blockIdx.x = b; emul_block_barrier = &bar;
        fused_tile<kMode, TH, NW>(v, &c, Qin, Qb, Qout, dt, src, mfx, ntx, *tile);
      });
    for (auto &x : th) x.join();
  }
}

extern "C" int axi_emul_steps(int W, int H, const double *xf, const double *rf_, const uint8_t *solid, const uint8_t *comp,
                              const float *zf /* 9 planes */, const double *comps, const double *fs, double cfl, int order, int visc,
                              double hf, float *Q /* 5N in/out */, float *src, float *mfx, int n, int fused) {
  const int N = W * H, NXF = (W + 1) * H, NRF = W * (H + 1);
  std::vector<float> rc(H), dx(W), dr(H), rf(H + 1);
  for (int i = 0; i < W; i++) dx[i] = (float)(xf[i + 1] - xf[i]);
  for (int j = 0; j < H; j++) { rc[j] = (float)(0.5 * (rf_[j] + rf_[j + 1])); dr[j] = (float)(rf_[j + 1] - rf_[j]); }
  for (int j = 0; j <= H; j++) rf[j] = (float)rf_[j];
  AxiView v{W, H, N, rc.data(), dx.data(), dr.data(), rf.data(), solid, comp, zf, zf + N, zf + 2 * N, zf + 3 * N, zf + 4 * N,
            zf + 5 * N, zf + 6 * N, zf + 7 * N, zf + 8 * N};
  AxiConsts c;
  std::memset(&c, 0, sizeof c);
  for (int q = 0; q < 32; q++) c.comps[q] = (float)comps[q];
  c.fs_r = (float)fs[0]; c.fs_u = (float)fs[1]; c.fs_p = (float)fs[2];
  c.cfl = (float)cfl; c.order = order; c.visc = visc; c.hf = (float)hf;
  std::vector<float> Q1(5 * N), Fx(5 * NXF), Fr(5 * NRF), nutF(NRF), dt(N);
  float cap = 0;
  const int NF = NXF + NRF;
  for (int it = 0; it < n; it++) {
    if (fused == 1) {          // tiles of 63 x 3, 4 waves (and rows from two rounds of waves)
      fused_launch<0, 3, 2>(v, c, Q, Q, Q1.data(), dt.data(), src, mfx);
      fused_launch<1, 3, 2>(v, c, Q1.data(), Q, Q, dt.data(), src, mfx);
      continue;
    }
    if (fused == 2) {
      fused_launch<0, 7, 8>(v, c, Q, Q, Q1.data(), dt.data(), src, mfx);
      fused_launch<1, 7, 8>(v, c, Q1.data(), Q, Q, dt.data(), src, mfx);
      continue;
    }
    launch(NF, [&] { face_kernel(v, &c, Q, Fx.data(), Fr.data(), nutF.data()); });
    launch(N, [&] { cell_kernel(v, &c, Q, Q, Q1.data(), Fx.data(), Fr.data(), nutF.data(), dt.data(), src, mfx, 0, &cap); });
    launch(NF, [&] { face_kernel(v, &c, Q1.data(), Fx.data(), Fr.data(), nutF.data()); });
    launch(N, [&] { cell_kernel(v, &c, Q1.data(), Q, Q, Fx.data(), Fr.data(), nutF.data(), dt.data(), src, mfx, 1, &cap); });
  }
  return 0;
}
