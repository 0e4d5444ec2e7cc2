// CPU run of the tiled stage kernel (3d_aero_lab_amd/csrc/euler3d_tile.hpp)
// for indexing checks: every block of the launch grid runs as TX*TY OS
// threads sharing one TileSmem and a barrier. Test infrastructure only.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "../../3d_aero_lab_amd/csrc/euler3d_tile.hpp"

thread_local emul_dim3 threadIdx;
thread_local emul_dim3 blockIdx;
thread_local std::barrier<> *emul_block_barrier = nullptr;

using namespace aero3d;

template <int TX, int TY>
static void run_stage(const CaseView &cv, const StageConsts &kc, const float *qin, const float *bse, float *out, float *dtout,
                      size_t ns, float wa, float wb, int zc) {
  const int ntx = (cv.NX + TX - 1) / TX, nty = (cv.NY + TY - 1) / TY, nzc = (cv.NZ + zc - 1) / zc;
  for (int lb = 0; lb < ntx * nty * nzc; lb++) {
    const int bx = lb % ntx, by = (lb / ntx) % nty, bz = lb / (ntx * nty);
    const int k0 = bz * zc, k1 = std::min(cv.NZ, k0 + zc);
    auto sm = std::make_unique<TileSmem<TX, TY>>();
    std::memset(sm.get(), 0xff, sizeof(TileSmem<TX, TY>));   // NaN-fill: a read of an unwritten slot shows up
    std::barrier<> bar(TX * TY);
    std::vector<std::thread> th;
    for (int y = 0; y < TY; y++)
      for (int x = 0; x < TX; x++)
        th.emplace_back([&, x, y] {
          threadIdx.x = x; threadIdx.y = y; blockIdx.x = lb; emul_block_barrier = &bar;
          tile_stage<TX, TY>(cv, kc, qin, bse, out, dtout, ns, wa, wb, bx * TX, by * TY, k0, k1, *sm);
        });
    for (auto &t : th) t.join();
  }
}

extern "C" int emul_steps(int NX, int NY, int NZ, const float *dx, const float *dy, const float *dz, const uint8_t *t,
                          const float *phi, const float *ap, const double *Fw, const double *fixw, const double *fixq,
                          double cfl, int order, float *Q, int n, int ty, int zc) {
  CaseView cv{NX, NY, NZ, NX * NY * NZ, dx, dy, dz, t, phi, ap};
  StageConsts kc;
  for (int v = 0; v < 5; v++) kc.Fw[v] = (float)Fw[v];
  for (int g = 0; g < 3; g++)
    for (int v = 0; v < 5; v++) { kc.fixw[g][v] = (float)fixw[g * 5 + v]; kc.fixq[g][v] = (float)fixq[g * 5 + v]; }
  kc.cfl = (float)cfl; kc.order = order;
  const size_t N = cv.N;
  // exact-size heap buffers so AddressSanitizer sees any overrun
  std::vector<float> S1(5 * N), B(5 * N), dt(N);
  // the solver keeps fixed cells at their conserved fixed states (reset, write, set_fix)
  for (size_t q = 0; q < N; q++)
    if (t[q] >= 3)
      for (int c = 0; c < 5; c++) Q[c * N + q] = kc.fixq[t[q] - 3][c];
  for (int it = 0; it < n; it++) {
    if (ty == 4) {
      run_stage<64, 4>(cv, kc, Q, Q, S1.data(), nullptr, N, 0.f, 1.f, zc);
      run_stage<64, 4>(cv, kc, S1.data(), Q, B.data(), dt.data(), N, 0.5f, 0.5f, zc);
    } else if (ty == 32) {          // 32 x 8 tiles
      run_stage<32, 8>(cv, kc, Q, Q, S1.data(), nullptr, N, 0.f, 1.f, zc);
      run_stage<32, 8>(cv, kc, S1.data(), Q, B.data(), dt.data(), N, 0.5f, 0.5f, zc);
    } else {
      run_stage<64, 8>(cv, kc, Q, Q, S1.data(), nullptr, N, 0.f, 1.f, zc);
      run_stage<64, 8>(cv, kc, S1.data(), Q, B.data(), dt.data(), N, 0.5f, 0.5f, zc);
    }
    std::memcpy(Q, B.data(), sizeof(float) * 5 * N);
  }
  return 0;
}
