// Test-only CPU build of the engine field model's device functions
// (3d_aero_lab_amd/csrc/field.hpp), so that they can be checked against the
// reference's outputs (and run under AddressSanitizer) before a device run.
// Never part of the product build.
#define __host__
#define __device__
#include <cstring>
#include "field.hpp"

extern "C" int fm_check(const double *T, int n, char *why, int nwhy) {
  const char *w = fieldm::check_table(T, n);
  if (!w) return 0;
  std::strncpy(why, w, nwhy - 1);
  why[nwhy - 1] = 0;
  return -1;
}

extern "C" void fm_sample(const double *T, const double *x, const double *r, const double *th, int n, double *out) {
  for (int p = 0; p < n; p++) {
    const fieldm::Sample o = fieldm::sample(T, x[p], r[p], th[p]);
    double *d = out + 14 * (size_t)p;
    d[0] = o.region; d[1] = o.T; d[2] = o.P; d[3] = o.Tt; d[4] = o.Pt; d[5] = o.V; d[6] = o.Vx; d[7] = o.Vth;
    d[8] = o.M; d[9] = o.a; d[10] = o.rho; d[11] = o.g; d[12] = o.cp; d[13] = o.R;
  }
}

extern "C" void fm_initial(const double *T, int W, int H, const double *xf, const double *rf, const uint8_t *solid,
                           const int8_t *stream, int uniform, double *Q) {
  const int N = W * H;
  for (int k = 0; k < N; k++) {
    const int i = k % W, j = k / W;
    const double x = 0.5 * (xf[i] + xf[i + 1]), r = 0.5 * (rf[j] + rf[j + 1]);
    const fieldm::Cons c = fieldm::initial_cell(T, x, r, solid[k], stream[k], uniform);
    for (int q = 0; q < 5; q++) Q[(size_t)q * N + k] = c.q[q];
  }
}
