// ============================================================================
// field.hip — C ABI of the engine field model on the device (axi_field_*,
// axi_initial_state; declared in axi.h): the FieldModel object of the
// reference (src/field.js:255-488) as a handle holding its table in device
// memory, point sampling, and the solver's initial state (initialState,
// src/cfd_grid.js:279-305) computed one thread per cell.
// ============================================================================
#include <hip/hip_runtime.h>

#include <string>

#include "axi.h"
#include "field.hpp"

struct axi_field {
  double *tab = nullptr;
  int ntab = 0, nstream = 0;
  hipStream_t stream = nullptr;
  std::string err;
};

namespace {

thread_local std::string g_field_err;
constexpr int kFB = 256;
constexpr int kMaxPoints = 1 << 26;

int ffail(axi_field *f, int code, const std::string &m) { if (f) f->err = m; else g_field_err = m; return code; }
#define FCHK(f, expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return ffail((f), -1000 - (int)e_, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// one thread per point: FieldModel.sample -> 14 doubles
__global__ void __launch_bounds__(kFB) field_sample_kernel(const double *__restrict__ T, const double *__restrict__ x,
                                                           const double *__restrict__ r, const double *__restrict__ th,
                                                           int n, double *__restrict__ out) {
  const int p = blockIdx.x * kFB + threadIdx.x;
  if (p >= n) return;
  const fieldm::Sample o = fieldm::sample(T, x[p], r[p], th[p]);
  double *d = out + 14 * (size_t)p;
  d[0] = o.region; d[1] = o.T; d[2] = o.P; d[3] = o.Tt; d[4] = o.Pt; d[5] = o.V; d[6] = o.Vx; d[7] = o.Vth;
  d[8] = o.M; d[9] = o.a; d[10] = o.rho; d[11] = o.g; d[12] = o.cp; d[13] = o.R;
}

// one thread per cell (row j, column i at j W + i): initialState
__global__ void __launch_bounds__(kFB) initial_state_kernel(const double *__restrict__ T, int W, int H,
                                                            const double *__restrict__ xf, const double *__restrict__ rf,
                                                            const uint8_t *__restrict__ solid, const int8_t *__restrict__ stream,
                                                            int uniform, double *__restrict__ Q) {
#ifdef __clang__
#pragma clang fp contract(off)
#endif
  const int k = blockIdx.x * kFB + threadIdx.x;
  const int N = W * H;
  if (k >= N) return;
  const int i = k % W, j = k / W;
  const double x = 0.5 * (xf[i] + xf[i + 1]), r = 0.5 * (rf[j] + rf[j + 1]);   // buildCFDGrid's centres
  const fieldm::Cons c = fieldm::initial_cell(T, x, r, solid[k], stream[k], uniform);
  for (int q = 0; q < 5; q++) Q[(size_t)q * N + k] = c.q[q];
}

template <class T>
int dalloc(axi_field *f, T **p, size_t n) {
  FCHK(f, hipMalloc((void **)p, sizeof(T) * (n ? n : 1)));
  return 0;
}

}  // namespace

extern "C" {

int axi_field_create(const double *tab, int ntab, axi_field **out) {
  if (!out) return ffail(nullptr, -1, "axi_field_create: null output");
  *out = nullptr;
  if (const char *why = fieldm::check_table(tab, ntab)) return ffail(nullptr, -1, why);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ffail(nullptr, -2, "axi_field_create: no HIP device");
  axi_field *f = new axi_field();
  f->ntab = ntab;
  f->nstream = (int)tab[AXI_FM_NSTREAM];
  int rc = 0;
  if (hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking) != hipSuccess) rc = ffail(nullptr, -2, "axi_field_create: stream");
  else if (hipMalloc((void **)&f->tab, sizeof(double) * ntab) != hipSuccess) rc = ffail(nullptr, -1002, "axi_field_create: hipMalloc failed");
  else if (hipMemcpy(f->tab, tab, sizeof(double) * ntab, hipMemcpyHostToDevice) != hipSuccess) rc = ffail(nullptr, -1001, "axi_field_create: upload failed");
  if (rc) { axi_field_destroy(f); return rc; }
  *out = f;
  return 0;
}

int axi_field_destroy(axi_field *f) {
  if (!f) return 0;
  if (f->tab) (void)hipFree(f->tab);
  if (f->stream) (void)hipStreamDestroy(f->stream);
  delete f;
  return 0;
}

const char *axi_field_last_error(const axi_field *f) { return f ? f->err.c_str() : g_field_err.c_str(); }

int axi_field_sample(axi_field *f, const double *x, const double *r, const double *th, int n, double *out) {
  if (!f) return ffail(nullptr, -1, "axi_field_sample: null handle");
  if (n < 0 || n > kMaxPoints || (n > 0 && (!x || !r || !th || !out))) return ffail(f, -1, "axi_field_sample: 0 <= n <= 2^26 and non-null arrays");
  if (n == 0) return 0;
  double *dx = nullptr, *dr = nullptr, *dth = nullptr, *dout = nullptr;
  int rc = dalloc(f, &dx, n);
  if (!rc) rc = dalloc(f, &dr, n);
  if (!rc) rc = dalloc(f, &dth, n);
  if (!rc) rc = dalloc(f, &dout, 14 * (size_t)n);
  auto run = [&]() -> int {
    FCHK(f, hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, f->stream));
    FCHK(f, hipMemcpyAsync(dr, r, sizeof(double) * n, hipMemcpyHostToDevice, f->stream));
    FCHK(f, hipMemcpyAsync(dth, th, sizeof(double) * n, hipMemcpyHostToDevice, f->stream));
    field_sample_kernel<<<(n + kFB - 1) / kFB, kFB, 0, f->stream>>>(f->tab, dx, dr, dth, n, dout);
    FCHK(f, hipGetLastError());
    FCHK(f, hipMemcpyAsync(out, dout, sizeof(double) * 14 * (size_t)n, hipMemcpyDeviceToHost, f->stream));
    FCHK(f, hipStreamSynchronize(f->stream));
    return 0;
  };
  if (!rc) rc = run();
  (void)hipFree(dx); (void)hipFree(dr); (void)hipFree(dth); (void)hipFree(dout);
  return rc;
}

int axi_initial_state(axi_field *f, int W, int H, const double *xf, const double *rf, const uint8_t *solid,
                      const int8_t *stream, int uniform, double *Q) {
  if (!f) return ffail(nullptr, -1, "axi_initial_state: null handle");
  if (W < 1 || H < 1 || (long long)W * H > (1LL << 28)) return ffail(f, -1, "axi_initial_state: W, H >= 1, W*H <= 2^28");
  if (!xf || !rf || !solid || !stream || !Q) return ffail(f, -1, "axi_initial_state: null array");
  const size_t N = (size_t)W * H;
  for (size_t k = 0; k < N; k++)
    if (stream[k] >= f->nstream) return ffail(f, -1, "axi_initial_state: zone stream index beyond the model's streams");
  double *dxf = nullptr, *drf = nullptr, *dQ = nullptr;
  uint8_t *dsolid = nullptr;
  int8_t *dstream = nullptr;
  int rc = dalloc(f, &dxf, W + 1);
  if (!rc) rc = dalloc(f, &drf, H + 1);
  if (!rc) rc = dalloc(f, &dsolid, N);
  if (!rc) rc = dalloc(f, &dstream, N);
  if (!rc) rc = dalloc(f, &dQ, 5 * N);
  auto run = [&]() -> int {
    FCHK(f, hipMemcpyAsync(dxf, xf, sizeof(double) * (W + 1), hipMemcpyHostToDevice, f->stream));
    FCHK(f, hipMemcpyAsync(drf, rf, sizeof(double) * (H + 1), hipMemcpyHostToDevice, f->stream));
    FCHK(f, hipMemcpyAsync(dsolid, solid, N, hipMemcpyHostToDevice, f->stream));
    FCHK(f, hipMemcpyAsync(dstream, stream, N, hipMemcpyHostToDevice, f->stream));
    initial_state_kernel<<<(unsigned)((N + kFB - 1) / kFB), kFB, 0, f->stream>>>(f->tab, W, H, dxf, drf, dsolid, dstream, uniform ? 1 : 0, dQ);
    FCHK(f, hipGetLastError());
    FCHK(f, hipMemcpyAsync(Q, dQ, sizeof(double) * 5 * N, hipMemcpyDeviceToHost, f->stream));
    FCHK(f, hipStreamSynchronize(f->stream));
    return 0;
  };
  if (!rc) rc = run();
  (void)hipFree(dxf); (void)hipFree(drf); (void)hipFree(dsolid); (void)hipFree(dstream); (void)hipFree(dQ);
  return rc;
}

}  // extern "C"
