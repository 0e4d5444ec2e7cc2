// ============================================================================
// aero3d_napi.cc — Node-API addon (aero3d.node) over the C ABI of
// libaero3d.so. Plain Node-API (N-API 8, no node-addon-api), typed arrays
// passed by pointer, every non-zero return code thrown as a JS Error with the
// library's message. The JS solver object CfdHIP (js/hip_solver.js) wraps it.
// ============================================================================
#include <node_api.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "aero3d.h"
#include "axi.h"

namespace {

#define NAPI_OK(call)                                                        \
  do {                                                                       \
    if ((call) != napi_ok) {                                                 \
      napi_throw_error(env, nullptr, "aero3d addon: Node-API call failed");  \
      return nullptr;                                                        \
    }                                                                        \
  } while (0)

napi_value throw_rc(napi_env env, aero3d_solver *s, int rc) {
  std::string msg = "libaero3d error " + std::to_string(rc) + ": " + aero3d_last_error(s);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

bool args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
    napi_throw_type_error(env, nullptr, "aero3d addon: missing arguments");
    return false;
  }
  return true;
}

// Typed-array data pointer and element count; type-checked.
template <class T>
bool typed(napi_env env, napi_value v, napi_typedarray_type want, T **data, size_t *n) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  napi_typedarray_type ty;
  void *p = nullptr;
  napi_value ab;
  size_t off = 0, len = 0;
  if (!is || napi_get_typedarray_info(env, v, &ty, &len, &p, &ab, &off) != napi_ok || ty != want) {
    napi_throw_type_error(env, nullptr, "aero3d addon: wrong typed-array type");
    return false;
  }
  *data = static_cast<T *>(p);
  if (n) *n = len;
  return true;
}

// A JS handle is an external pointing at one of these records: the solver and
// its cell count (every array argument is checked against it). The addon keeps
// the set of live records, so a foreign external, a handle of the other solver
// kind or a handle used after dispose is a JS error, not a stray pointer. The
// external's finalizer frees the record and, for a handle never disposed, the
// solver and its device memory.
struct Rec3 { aero3d_solver *s; size_t N; bool busy = false; };
struct RecAxi { axi_solver *s; size_t N; int W, H; };
struct RecField { axi_field *s; };
std::mutex g_mu;
std::unordered_set<const void *> g_rec3, g_recaxi, g_recfield;

void fin3(napi_env, void *d, void *) {
  Rec3 *r = static_cast<Rec3 *>(d);
  { std::lock_guard<std::mutex> l(g_mu); g_rec3.erase(r); }
  if (r->s) aero3d_destroy(r->s);
  delete r;
}
void finaxi(napi_env, void *d, void *) {
  RecAxi *r = static_cast<RecAxi *>(d);
  { std::lock_guard<std::mutex> l(g_mu); g_recaxi.erase(r); }
  if (r->s) axi_destroy(r->s);
  delete r;
}
void finfield(napi_env, void *d, void *) {
  RecField *r = static_cast<RecField *>(d);
  { std::lock_guard<std::mutex> l(g_mu); g_recfield.erase(r); }
  if (r->s) axi_field_destroy(r->s);
  delete r;
}

// A handle with asynchronous work in flight may not be used (the library's
// handles are single-threaded); the axisymmetric solver has no async calls.
template <class R> bool busy_of(const R *r) { return r->busy; }
template <> bool busy_of<RecAxi>(const RecAxi *) { return false; }
template <> bool busy_of<RecField>(const RecField *) { return false; }

template <class R>
R *record(napi_env env, napi_value v, const std::unordered_set<const void *> &live, const char *what) {
  void *p = nullptr;
  bool ok = napi_get_value_external(env, v, &p) == napi_ok && p;
  if (ok) { std::lock_guard<std::mutex> l(g_mu); ok = live.count(p) > 0; }
  if (!ok) {
    std::string m = std::string("aero3d addon: not ") + what;
    napi_throw_type_error(env, nullptr, m.c_str());
    return nullptr;
  }
  R *r = static_cast<R *>(p);
  if (!r->s) { napi_throw_error(env, nullptr, "aero3d addon: solver already disposed"); return nullptr; }
  if (busy_of(r)) { napi_throw_error(env, nullptr, "aero3d addon: handle busy with an asynchronous call"); return nullptr; }
  return r;
}
Rec3 *rec3(napi_env env, napi_value v) { return record<Rec3>(env, v, g_rec3, "a 3D solver handle"); }

aero3d_solver *handle(napi_env env, napi_value v) {
  Rec3 *r = rec3(env, v);
  return r ? r->s : nullptr;
}

bool check_len(napi_env env, size_t n, size_t want, const char *what) {
  if (n == want) return true;
  std::string m = std::string("aero3d addon: ") + what + " has " + std::to_string(n) + " elements, needs " + std::to_string(want);
  napi_throw_range_error(env, nullptr, m.c_str());
  return false;
}

double num(napi_env env, napi_value v) { double d = 0; napi_get_value_double(env, v, &d); return d; }

napi_value undef(napi_env env) { napi_value u; napi_get_undefined(env, &u); return u; }

// create(fx, fy, fz, t, phi, ap0, ap1, ap2, ap3) -> external handle
napi_value Create(napi_env env, napi_callback_info info) {
  napi_value a[9];
  if (!args(env, info, 9, a)) return nullptr;
  double *fx, *fy, *fz; size_t nx, ny, nz, n, nphi, nap[4];
  uint8_t *t; float *phi, *ap[4];
  if (!typed(env, a[0], napi_float64_array, &fx, &nx) || !typed(env, a[1], napi_float64_array, &fy, &ny) ||
      !typed(env, a[2], napi_float64_array, &fz, &nz) || !typed(env, a[3], napi_uint8_array, &t, &n) ||
      !typed(env, a[4], napi_float32_array, &phi, &nphi))
    return nullptr;
  for (int d = 0; d < 4; d++) if (!typed(env, a[5 + d], napi_float32_array, &ap[d], &nap[d])) return nullptr;
  if (nx < 2 || ny < 2 || nz < 2 || nx > (1u << 30) || ny > (1u << 30) || nz > (1u << 30)) {
    napi_throw_range_error(env, nullptr, "aero3d addon: each axis needs at least two faces");
    return nullptr;
  }
  const size_t N = (nx - 1) * (ny - 1) * (nz - 1);
  if (!check_len(env, n, N, "t") || !check_len(env, nphi, N, "phi")) return nullptr;
  for (int d = 0; d < 4; d++) if (!check_len(env, nap[d], N, "ap")) return nullptr;
  aero3d_grid g{(int)nx - 1, (int)ny - 1, (int)nz - 1, fx, fy, fz};
  aero3d_solver *s = nullptr;
  int rc = aero3d_create(&g, t, phi, ap[0], ap[1], ap[2], ap[3], &s);
  if (rc) return throw_rc(env, nullptr, rc);
  Rec3 *r = new Rec3{s, N};
  { std::lock_guard<std::mutex> l(g_mu); g_rec3.insert(r); }
  napi_value ext;
  if (napi_create_external(env, r, fin3, nullptr, &ext) != napi_ok) {
    fin3(env, r, nullptr);
    napi_throw_error(env, nullptr, "aero3d addon: Node-API call failed");
    return nullptr;
  }
  return ext;
}

napi_value Dispose(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  Rec3 *r = rec3(env, a[0]);
  if (!r) return nullptr;
  aero3d_destroy(r->s);
  r->s = nullptr;                // the record lives until the external is collected
  return undef(env);
}

napi_value SetFree(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  double *F; size_t n;
  if (!s || !typed(env, a[1], napi_float64_array, &F, &n)) return nullptr;
  if (n != 5) { napi_throw_range_error(env, nullptr, "setFree: 5 values"); return nullptr; }
  int rc = aero3d_set_free(s, F);
  return rc ? throw_rc(env, s, rc) : undef(env);
}

napi_value SetFix(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  double *F; size_t n;
  if (!s || !typed(env, a[1], napi_float64_array, &F, &n)) return nullptr;
  if (n != 15) { napi_throw_range_error(env, nullptr, "setFix: 3 x 5 values"); return nullptr; }
  int rc = aero3d_set_fix(s, F);
  return rc ? throw_rc(env, s, rc) : undef(env);
}

napi_value SetOpt(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!args(env, info, 3, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  int rc = aero3d_set_opt(s, num(env, a[1]), (int)num(env, a[2]));
  return rc ? throw_rc(env, s, rc) : undef(env);
}

napi_value Reset(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  int rc = aero3d_reset(s);
  return rc ? throw_rc(env, s, rc) : undef(env);
}

napi_value Step(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  int rc = aero3d_step(s, (int)num(env, a[1]));
  return rc ? throw_rc(env, s, rc) : undef(env);
}

napi_value Iter(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  int it = 0;
  int rc = aero3d_iter(s, &it);
  if (rc) return throw_rc(env, s, rc);
  napi_value v;
  NAPI_OK(napi_create_int32(env, it, &v));
  return v;
}

// read(h, r, ru, rv, rw, E, dt|null): fills caller Float32Arrays
napi_value Read(napi_env env, napi_callback_info info) {
  napi_value a[7];
  if (!args(env, info, 7, a)) return nullptr;
  Rec3 *r = rec3(env, a[0]);
  if (!r) return nullptr;
  aero3d_solver *s = r->s;
  float *p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int c = 0; c < 6; c++) {
    napi_valuetype vt;
    napi_typeof(env, a[1 + c], &vt);
    if (vt == napi_null || vt == napi_undefined) continue;
    size_t n;
    if (!typed(env, a[1 + c], napi_float32_array, &p[c], &n) || !check_len(env, n, r->N, "read output")) return nullptr;
  }
  int rc = aero3d_read(s, p[0], p[1], p[2], p[3], p[4], p[5]);
  return rc ? throw_rc(env, s, rc) : undef(env);
}

napi_value Write(napi_env env, napi_callback_info info) {
  napi_value a[6];
  if (!args(env, info, 6, a)) return nullptr;
  Rec3 *r = rec3(env, a[0]);
  if (!r) return nullptr;
  aero3d_solver *s = r->s;
  float *p[5];
  for (int c = 0; c < 5; c++) {
    size_t n;
    if (!typed(env, a[1 + c], napi_float32_array, &p[c], &n) || !check_len(env, n, r->N, "write input")) return nullptr;
  }
  int rc = aero3d_write(s, p[0], p[1], p[2], p[3], p[4]);
  return rc ? throw_rc(env, s, rc) : undef(env);
}

napi_value Sync(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  int rc = aero3d_sync(s);
  return rc ? throw_rc(env, s, rc) : undef(env);
}

// setForceList(h, q Int32Array, cx, cz, x, z Float64Array)
napi_value SetForceList(napi_env env, napi_callback_info info) {
  napi_value a[6];
  if (!args(env, info, 6, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  int32_t *q; size_t n; double *d[4];
  if (!s || !typed(env, a[1], napi_int32_array, &q, &n)) return nullptr;
  for (int c = 0; c < 4; c++) {
    size_t m;
    if (!typed(env, a[2 + c], napi_float64_array, &d[c], &m) || !check_len(env, m, n, "force-list column")) return nullptr;
  }
  int rc = aero3d_set_force_list(s, q, d[0], d[1], d[2], d[3], (int)n);
  return rc ? throw_rc(env, s, rc) : undef(env);
}

// forces(h, alpha, M, S, xcg, zcg, cref) -> {CL, CD, Cm}
napi_value Forces(napi_env env, napi_callback_info info) {
  napi_value a[7];
  if (!args(env, info, 7, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  double out[3];
  int rc = aero3d_forces(s, num(env, a[1]), num(env, a[2]), num(env, a[3]), num(env, a[4]), num(env, a[5]), num(env, a[6]), out);
  if (rc) return throw_rc(env, s, rc);
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  const char *k[3] = {"CL", "CD", "Cm"};
  for (int c = 0; c < 3; c++) { NAPI_OK(napi_create_double(env, out[c], &v)); NAPI_OK(napi_set_named_property(env, o, k[c], v)); }
  return o;
}

napi_value Trefftz(napi_env env, napi_callback_info info) {
  napi_value a[5];
  if (!args(env, info, 5, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  double cdi;
  int rc = aero3d_trefftz(s, num(env, a[1]), num(env, a[2]), num(env, a[3]), num(env, a[4]), &cdi);
  if (rc) return throw_rc(env, s, rc);
  napi_value v;
  NAPI_OK(napi_create_double(env, cdi, &v));
  return v;
}

napi_value Residual(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  double r;
  int rc = aero3d_residual(s, &r);
  if (rc) return throw_rc(env, s, rc);
  napi_value v;
  NAPI_OK(napi_create_double(env, r, &v));
  return v;
}

napi_value TimeSteps(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  float ms;
  int rc = aero3d_time_steps(s, (int)num(env, a[1]), &ms);
  if (rc) return throw_rc(env, s, rc);
  napi_value v;
  NAPI_OK(napi_create_double(env, ms, &v));
  return v;
}

napi_value Autotune(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  int ty, zc;
  float ms;
  int rc = aero3d_autotune(s, (int)num(env, a[1]), &ty, &zc, &ms);
  if (rc) return throw_rc(env, s, rc);
  napi_value o, v;
  NAPI_OK(napi_create_object(env, &o));
  NAPI_OK(napi_create_int32(env, ty, &v)); NAPI_OK(napi_set_named_property(env, o, "ty", v));
  NAPI_OK(napi_create_int32(env, zc, &v)); NAPI_OK(napi_set_named_property(env, o, "zchunk", v));
  NAPI_OK(napi_create_double(env, ms, &v)); NAPI_OK(napi_set_named_property(env, o, "msPerIter", v));
  return o;
}

void set_num(napi_env env, napi_value o, const char *k, double x) {
  napi_value v;
  if (napi_create_double(env, x, &v) == napi_ok) napi_set_named_property(env, o, k, v);
}

// readback(h, alpha, M, S, xcg, zcg, cref, xT) -> {it, res, CL, CD, Cm, CDi, Mmax}
napi_value Readback(napi_env env, napi_callback_info info) {
  napi_value a[8];
  if (!args(env, info, 8, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  aero3d_readback_in in{num(env, a[1]), num(env, a[2]), num(env, a[3]), num(env, a[4]), num(env, a[5]), num(env, a[6]), num(env, a[7])};
  aero3d_readback_out out;
  int rc = aero3d_readback(s, &in, &out);
  if (rc) return throw_rc(env, s, rc);
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  set_num(env, o, "it", out.it); set_num(env, o, "res", out.res); set_num(env, o, "CL", out.CL); set_num(env, o, "CD", out.CD);
  set_num(env, o, "Cm", out.Cm); set_num(env, o, "CDi", out.CDi); set_num(env, o, "Mmax", out.Mmax);
  return o;
}

// optional typed array: null / undefined give nullptr; otherwise type- and length-checked
template <class T>
bool opt_typed(napi_env env, napi_value v, napi_typedarray_type want, size_t len, T **data, const char *what) {
  napi_valuetype ty;
  if (napi_typeof(env, v, &ty) != napi_ok) return false;
  if (ty == napi_null || ty == napi_undefined) { *data = nullptr; return true; }
  size_t n;
  return typed(env, v, want, data, &n) && check_len(env, n, len, what);
}

// probe(h, pts Float64Array(3n), fs Float64Array(5) [T, a, P, rho, q] | null, M,
//       idx Int32Array(8n), w Float32Array(8n), ok Uint8Array(n), W Float64Array(5n),
//       state Float64Array(12n), vel Float64Array(3n), vok Uint8Array(n))   (outputs may be null)
napi_value Probe(napi_env env, napi_callback_info info) {
  napi_value a[11];
  if (!args(env, info, 11, a)) return nullptr;
  aero3d_solver *s = handle(env, a[0]);
  if (!s) return nullptr;
  double *pts;
  size_t np;
  if (!typed(env, a[1], napi_float64_array, &pts, &np)) return nullptr;
  if (np % 3) { napi_throw_range_error(env, nullptr, "aero3d addon: probe points need 3 coordinates each"); return nullptr; }
  const size_t n = np / 3;
  double *fsv;
  if (!opt_typed(env, a[2], napi_float64_array, 5, &fsv, "probe freestream record")) return nullptr;
  aero3d_probe_out o{};
  if (!opt_typed(env, a[4], napi_int32_array, 8 * n, &o.idx, "probe idx") ||
      !opt_typed(env, a[5], napi_float32_array, 8 * n, &o.w, "probe w") ||
      !opt_typed(env, a[6], napi_uint8_array, n, &o.ok, "probe ok") ||
      !opt_typed(env, a[7], napi_float64_array, 5 * n, &o.W, "probe W") ||
      !opt_typed(env, a[8], napi_float64_array, 12 * n, &o.state, "probe state") ||
      !opt_typed(env, a[9], napi_float64_array, 3 * n, &o.vel, "probe vel") ||
      !opt_typed(env, a[10], napi_uint8_array, n, &o.vok, "probe vok"))
    return nullptr;
  aero3d_probe_fs fs{};
  if (fsv) fs = {fsv[0], fsv[1], fsv[2], fsv[3], fsv[4]};
  int rc = aero3d_probe(s, (int)std::min(n, (size_t)1 << 30), pts, fsv ? &fs : nullptr, num(env, a[3]), &o);
  return rc ? throw_rc(env, s, rc) : undef(env);
}

// ---------------------------------------------------------------------------
// asynchronous calls: the library call runs on a worker thread
// (napi_async_work) and settles a Promise; the handle is busy meanwhile and a
// reference keeps it alive.
// ---------------------------------------------------------------------------
struct AsyncCall {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  napi_ref keep = nullptr;           // the JS handle, alive until settled
  bool *busy = nullptr;
  std::function<int(std::string &)> run;           // worker thread: the library call (rc, message)
  std::function<napi_value(napi_env)> result;      // main thread: the resolved value
  int rc = 0;
  std::string msg;
};
void async_execute(napi_env, void *d) {
  AsyncCall *c = static_cast<AsyncCall *>(d);
  c->rc = c->run(c->msg);
}
void async_complete(napi_env env, napi_status, void *d) {
  AsyncCall *c = static_cast<AsyncCall *>(d);
  *c->busy = false;
  if (c->rc) {
    napi_value m, e;
    napi_create_string_utf8(env, c->msg.c_str(), NAPI_AUTO_LENGTH, &m);
    napi_create_error(env, nullptr, m, &e);
    napi_reject_deferred(env, c->deferred, e);
  } else {
    napi_value v = c->result ? c->result(env) : nullptr;
    if (!v) napi_get_undefined(env, &v);
    napi_resolve_deferred(env, c->deferred, v);
  }
  napi_delete_reference(env, c->keep);
  napi_delete_async_work(env, c->work);
  delete c;
}
napi_value start_async(napi_env env, napi_value handle_value, bool *busy, AsyncCall *c) {
  napi_value promise, name;
  c->busy = busy;
  if (napi_create_promise(env, &c->deferred, &promise) != napi_ok || napi_create_reference(env, handle_value, 1, &c->keep) != napi_ok ||
      napi_create_string_utf8(env, "aero3d", NAPI_AUTO_LENGTH, &name) != napi_ok ||
      napi_create_async_work(env, nullptr, name, async_execute, async_complete, c, &c->work) != napi_ok ||
      napi_queue_async_work(env, c->work) != napi_ok) {
    delete c;
    napi_throw_error(env, nullptr, "aero3d addon: cannot start asynchronous work");
    return nullptr;
  }
  *busy = true;
  return promise;
}

// stepAsync(h, n) -> Promise: n iterations, settled once they have run on the device
napi_value StepAsync(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  Rec3 *r = rec3(env, a[0]);
  if (!r) return nullptr;
  aero3d_solver *s = r->s;
  const int n = (int)num(env, a[1]);
  AsyncCall *c = new AsyncCall();
  c->run = [s, n](std::string &m) {
    int rc = aero3d_step(s, n);
    if (!rc) rc = aero3d_sync(s);
    if (rc) m = "libaero3d error " + std::to_string(rc) + ": " + aero3d_last_error(s);
    return rc;
  };
  return start_async(env, a[0], &r->busy, c);
}

// ---------------------------------------------------------------------------
// batched form (aero3d_batch_*) and served form (aero3d_server_*)
// ---------------------------------------------------------------------------
struct RecB { aero3d_batch *b; bool busy = false; std::unordered_map<int, size_t> ncell; };   // cells per geometry and per slot (key -1 - slot)
struct RecS { aero3d_server *s; RecB batch; };
std::unordered_set<const void *> g_recb, g_recs;

void finb(napi_env, void *d, void *) {
  RecB *r = static_cast<RecB *>(d);
  { std::lock_guard<std::mutex> l(g_mu); g_recb.erase(r); }
  if (r->b) aero3d_batch_destroy(r->b);
  delete r;
}
void fins(napi_env, void *d, void *) {
  RecS *r = static_cast<RecS *>(d);
  { std::lock_guard<std::mutex> l(g_mu); g_recs.erase(r); }
  if (r->s) aero3d_server_destroy(r->s);
  delete r;
}
template <> bool busy_of<RecS>(const RecS *r) { return r->batch.busy; }
RecB *recb(napi_env env, napi_value v) {
  void *p = nullptr;
  bool ok = napi_get_value_external(env, v, &p) == napi_ok && p;
  if (ok) { std::lock_guard<std::mutex> l(g_mu); ok = g_recb.count(p) > 0; }
  if (!ok) { napi_throw_type_error(env, nullptr, "aero3d addon: not a batch handle"); return nullptr; }
  RecB *r = static_cast<RecB *>(p);
  if (!r->b) { napi_throw_error(env, nullptr, "aero3d addon: batch already disposed"); return nullptr; }
  if (r->busy) { napi_throw_error(env, nullptr, "aero3d addon: handle busy with an asynchronous call"); return nullptr; }
  return r;
}
RecS *recs(napi_env env, napi_value v) { return record<RecS>(env, v, g_recs, "a server handle"); }
napi_value batch_throw(napi_env env, aero3d_batch *b, int rc) {
  std::string msg = "libaero3d batch error " + std::to_string(rc) + ": " + aero3d_batch_last_error(b);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}
template <class R, class F>
napi_value make_external(napi_env env, R *r, std::unordered_set<const void *> &live, F fin) {
  { std::lock_guard<std::mutex> l(g_mu); live.insert(r); }
  napi_value ext;
  if (napi_create_external(env, r, fin, nullptr, &ext) != napi_ok) {
    fin(env, r, nullptr);
    napi_throw_error(env, nullptr, "aero3d addon: Node-API call failed");
    return nullptr;
  }
  return ext;
}
int64_t i64(napi_env env, napi_value v) { int64_t x = 0; napi_get_value_int64(env, v, &x); return x; }

// batchCreate(arenaBytes, maxSlots) -> handle
napi_value BatchCreate(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  aero3d_batch *b = nullptr;
  int rc = aero3d_batch_create(i64(env, a[0]), (int)num(env, a[1]), &b);
  if (rc) return batch_throw(env, nullptr, rc);
  return make_external(env, new RecB{b}, g_recb, finb);
}
napi_value BatchDispose(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  aero3d_batch_destroy(r->b);
  r->b = nullptr;
  return undef(env);
}
// geometry arguments a[0..13]: fx, fy, fz, t, phi, ap0..ap3, fq, fcx, fcz, fxc, fzc
int add_geometry(napi_env env, aero3d_batch *b, RecB *rb, napi_value *a, int *gid) {
  double *f[3]; size_t nf3[3];
  for (int d = 0; d < 3; d++) if (!typed(env, a[d], napi_float64_array, &f[d], &nf3[d])) return -1;
  if (nf3[0] < 2 || nf3[1] < 2 || nf3[2] < 2 || nf3[0] > (1u << 30) || nf3[1] > (1u << 30) || nf3[2] > (1u << 30)) {
    napi_throw_range_error(env, nullptr, "addGeometry: each axis needs at least two faces");
    return -1;
  }
  const size_t N = (nf3[0] - 1) * (nf3[1] - 1) * (nf3[2] - 1);
  uint8_t *t; float *phi, *ap[4]; size_t n;
  if (!typed(env, a[3], napi_uint8_array, &t, &n) || !check_len(env, n, N, "t")) return -1;
  if (!typed(env, a[4], napi_float32_array, &phi, &n) || !check_len(env, n, N, "phi")) return -1;
  for (int d = 0; d < 4; d++) if (!typed(env, a[5 + d], napi_float32_array, &ap[d], &n) || !check_len(env, n, N, "ap")) return -1;
  int32_t *fq; size_t nfl; double *fd[4];
  if (!typed(env, a[9], napi_int32_array, &fq, &nfl)) return -1;
  for (int c = 0; c < 4; c++) if (!typed(env, a[10 + c], napi_float64_array, &fd[c], &n) || !check_len(env, n, nfl, "force-list column")) return -1;
  aero3d_grid g{(int)nf3[0] - 1, (int)nf3[1] - 1, (int)nf3[2] - 1, f[0], f[1], f[2]};
  int rc = aero3d_batch_add_geometry(b, &g, t, phi, ap[0], ap[1], ap[2], ap[3], fq, fd[0], fd[1], fd[2], fd[3], (int)nfl, gid);
  if (rc) { batch_throw(env, b, rc); return -1; }
  rb->ncell[*gid] = N;
  return 0;
}
napi_value int_value(napi_env env, int x) { napi_value v; napi_create_int32(env, x, &v); return v; }
// batchAddGeometry(h, fx, fy, fz, t, phi, ap0, ap1, ap2, ap3, fq, fcx, fcz, fxc, fzc) -> gid
napi_value BatchAddGeometry(napi_env env, napi_callback_info info) {
  napi_value a[15];
  if (!args(env, info, 15, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  int gid;
  if (!r || add_geometry(env, r->b, r, a + 1, &gid)) return nullptr;
  return int_value(env, gid);
}
napi_value BatchRemoveGeometry(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  int rc = aero3d_batch_remove_geometry(r->b, (int)num(env, a[1]));
  return rc ? batch_throw(env, r->b, rc) : undef(env);
}
// batchAdd(h, gid, F Float64Array(5), fix Float64Array(15), cfl, order) -> slot
napi_value BatchAdd(napi_env env, napi_callback_info info) {
  napi_value a[6];
  if (!args(env, info, 6, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  double *F, *fix; size_t nF, nfix;
  if (!r || !typed(env, a[2], napi_float64_array, &F, &nF) || !typed(env, a[3], napi_float64_array, &fix, &nfix) ||
      !check_len(env, nF, 5, "F") || !check_len(env, nfix, 15, "fix"))
    return nullptr;
  const int gid = (int)num(env, a[1]);
  int slot;
  int rc = aero3d_batch_add(r->b, gid, F, fix, num(env, a[4]), (int)num(env, a[5]), &slot);
  if (rc) return batch_throw(env, r->b, rc);
  r->ncell[-1 - slot] = r->ncell[gid];
  return int_value(env, slot);
}
napi_value BatchRemove(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  int rc = aero3d_batch_remove(r->b, (int)num(env, a[1]));
  return rc ? batch_throw(env, r->b, rc) : undef(env);
}
napi_value BatchStep(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  int rc = aero3d_batch_step(r->b, (int)num(env, a[1]));
  return rc ? batch_throw(env, r->b, rc) : undef(env);
}
// batchStepAsync(h, n) -> Promise, settled once the iterations have run
napi_value BatchStepAsync(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  aero3d_batch *b = r->b;
  const int n = (int)num(env, a[1]);
  AsyncCall *c = new AsyncCall();
  c->run = [b, n](std::string &m) {
    int rc = aero3d_batch_step(b, n);
    if (!rc) rc = aero3d_batch_sync(b);
    if (rc) m = "libaero3d batch error " + std::to_string(rc) + ": " + aero3d_batch_last_error(b);
    return rc;
  };
  return start_async(env, a[0], &r->busy, c);
}
napi_value BatchSync(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  int rc = aero3d_batch_sync(r->b);
  return rc ? batch_throw(env, r->b, rc) : undef(env);
}
napi_value BatchIter(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  int it = 0;
  int rc = aero3d_batch_iter(r->b, (int)num(env, a[1]), &it);
  return rc ? batch_throw(env, r->b, rc) : int_value(env, it);
}
// batchRead(h, slot, r, ru, rv, rw, E): Float32Arrays of the case's cell count
napi_value BatchRead(napi_env env, napi_callback_info info) {
  napi_value a[7];
  if (!args(env, info, 7, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  const int slot = (int)num(env, a[1]);
  auto it = r->ncell.find(-1 - slot);
  if (it == r->ncell.end()) { napi_throw_range_error(env, nullptr, "batchRead: unknown slot"); return nullptr; }
  float *p[5];
  for (int c = 0; c < 5; c++) {
    size_t n;
    if (!typed(env, a[2 + c], napi_float32_array, &p[c], &n) || !check_len(env, n, it->second, "batchRead output")) return nullptr;
  }
  int rc = aero3d_batch_read(r->b, slot, p[0], p[1], p[2], p[3], p[4]);
  return rc ? batch_throw(env, r->b, rc) : undef(env);
}
// batchResiduals(h, slots Int32Array, out Float64Array)
napi_value BatchResiduals(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!args(env, info, 3, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  int32_t *sl; double *o; size_t n, m;
  if (!r || !typed(env, a[1], napi_int32_array, &sl, &n) || !typed(env, a[2], napi_float64_array, &o, &m) ||
      !check_len(env, m, n, "residual output"))
    return nullptr;
  int rc = aero3d_batch_residuals(r->b, sl, (int)n, o);
  return rc ? batch_throw(env, r->b, rc) : undef(env);
}
// batchForces(h, slot, alpha, M, S, xcg, zcg, cref) -> {CL, CD, Cm}
napi_value BatchForces(napi_env env, napi_callback_info info) {
  napi_value a[8];
  if (!args(env, info, 8, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  double out[3];
  int rc = aero3d_batch_forces(r->b, (int)num(env, a[1]), num(env, a[2]), num(env, a[3]), num(env, a[4]), num(env, a[5]), num(env, a[6]),
                               num(env, a[7]), out);
  if (rc) return batch_throw(env, r->b, rc);
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  set_num(env, o, "CL", out[0]); set_num(env, o, "CD", out[1]); set_num(env, o, "Cm", out[2]);
  return o;
}
// batchSetOption(h, name, value): zchunk | tileWidth | persistent | groupCells
napi_value BatchSetOption(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!args(env, info, 3, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  char key[32] = {0};
  size_t len = 0;
  napi_get_value_string_utf8(env, a[1], key, sizeof key, &len);
  const std::string k(key, len);
  int rc;
  if (k == "zchunk") rc = aero3d_batch_set_zchunk(r->b, (int)num(env, a[2]));
  else if (k == "tileWidth") rc = aero3d_batch_set_tile_width(r->b, (int)num(env, a[2]));
  else if (k == "persistent") rc = aero3d_batch_set_persistent(r->b, (int)num(env, a[2]));
  else if (k == "groupCells") rc = aero3d_batch_set_group_cells(r->b, i64(env, a[2]));
  else { napi_throw_range_error(env, nullptr, "batchSetOption: zchunk, tileWidth, persistent or groupCells"); return nullptr; }
  return rc ? batch_throw(env, r->b, rc) : undef(env);
}
napi_value BatchFootprint(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecB *r = recb(env, a[0]);
  if (!r) return nullptr;
  int64_t used, cap;
  int rc = aero3d_batch_footprint(r->b, &used, &cap);
  if (rc) return batch_throw(env, r->b, rc);
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  set_num(env, o, "used", (double)used); set_num(env, o, "capacity", (double)cap);
  return o;
}

// serverCreate(arenaBytes, maxSlots, checkEvery) -> handle
napi_value ServerCreate(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!args(env, info, 3, a)) return nullptr;
  aero3d_server *s = nullptr;
  int rc = aero3d_server_create(i64(env, a[0]), (int)num(env, a[1]), (int)num(env, a[2]), &s);
  if (rc) return batch_throw(env, nullptr, rc);
  RecS *r = new RecS{s, RecB{aero3d_server_batch(s)}};
  return make_external(env, r, g_recs, fins);
}
napi_value ServerDispose(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecS *r = recs(env, a[0]);
  if (!r) return nullptr;
  aero3d_server_destroy(r->s);
  r->s = nullptr;
  return undef(env);
}
// serverAddGeometry(h, <the 14 geometry arrays of batchAddGeometry>) -> gid
napi_value ServerAddGeometry(napi_env env, napi_callback_info info) {
  napi_value a[15];
  if (!args(env, info, 15, a)) return nullptr;
  RecS *r = recs(env, a[0]);
  int gid;
  if (!r || add_geometry(env, aero3d_server_batch(r->s), &r->batch, a + 1, &gid)) return nullptr;
  return int_value(env, gid);
}
napi_value ServerRemoveGeometry(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  RecS *r = recs(env, a[0]);
  if (!r) return nullptr;
  aero3d_batch *b = aero3d_server_batch(r->s);
  int rc = aero3d_batch_remove_geometry(b, (int)num(env, a[1]));
  return rc ? batch_throw(env, b, rc) : undef(env);
}
bool get_req(napi_env env, napi_value o, const char *k, napi_value *v) {
  bool has = false;
  if (napi_has_named_property(env, o, k, &has) != napi_ok || !has || napi_get_named_property(env, o, k, v) != napi_ok) {
    std::string m = std::string("serverSubmit: request.") + k + " missing";
    napi_throw_type_error(env, nullptr, m.c_str());
    return false;
  }
  return true;
}
double prop_num(napi_env env, napi_value o, const char *k, double dflt) {
  bool has = false;
  napi_value v;
  if (napi_has_named_property(env, o, k, &has) != napi_ok || !has || napi_get_named_property(env, o, k, &v) != napi_ok) return dflt;
  napi_valuetype t;
  napi_typeof(env, v, &t);
  return t == napi_number ? num(env, v) : dflt;
}
// serverSubmit(h, {geometry, F: Float64Array(5), fix: Float64Array(15), cfl, order, alpha, M, S, xcg, zcg, cref,
//                  rule, resDrop, maxIters, arrival, tag})
napi_value ServerSubmit(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  RecS *r = recs(env, a[0]);
  if (!r) return nullptr;
  aero3d_request q;
  std::memset(&q, 0, sizeof q);
  napi_value v;
  double *F, *fix; size_t n;
  if (!get_req(env, a[1], "F", &v) || !typed(env, v, napi_float64_array, &F, &n) || !check_len(env, n, 5, "F")) return nullptr;
  std::memcpy(q.F, F, sizeof q.F);
  if (!get_req(env, a[1], "fix", &v) || !typed(env, v, napi_float64_array, &fix, &n) || !check_len(env, n, 15, "fix")) return nullptr;
  std::memcpy(q.fix, fix, sizeof q.fix);
  q.geometry = (int)prop_num(env, a[1], "geometry", -1);
  q.cfl = prop_num(env, a[1], "cfl", 0.5); q.order = (int)prop_num(env, a[1], "order", 2);
  q.alpha = prop_num(env, a[1], "alpha", 0); q.M = prop_num(env, a[1], "M", 0); q.S = prop_num(env, a[1], "S", 1);
  q.xcg = prop_num(env, a[1], "xcg", 0); q.zcg = prop_num(env, a[1], "zcg", 0); q.cref = prop_num(env, a[1], "cref", 1);
  q.rule = (int)prop_num(env, a[1], "rule", AERO3D_RULE_RES_DROP); q.res_drop = prop_num(env, a[1], "resDrop", 1e-4);
  q.max_iters = (int)prop_num(env, a[1], "maxIters", 2000); q.arrival = prop_num(env, a[1], "arrival", 0);
  q.tag = (int64_t)prop_num(env, a[1], "tag", 0);
  q.state_out = nullptr;
  aero3d_batch *b = aero3d_server_batch(r->s);
  int rc = aero3d_server_submit(r->s, &q);
  return rc ? batch_throw(env, b, rc) : undef(env);
}
napi_value tick_value(napi_env env, int act, int pend) {
  napi_value o;
  if (napi_create_object(env, &o) != napi_ok) return nullptr;
  set_num(env, o, "active", act); set_num(env, o, "pending", pend);
  return o;
}
// serverTick(h) -> {active, pending}
napi_value ServerTick(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecS *r = recs(env, a[0]);
  if (!r) return nullptr;
  int act = 0, pend = 0;
  int rc = aero3d_server_tick(r->s, &act, &pend);
  return rc ? batch_throw(env, aero3d_server_batch(r->s), rc) : tick_value(env, act, pend);
}
// serverTickAsync(h) -> Promise<{active, pending}>: the tick (and its wait for
// the older tick's results) runs on a worker thread
napi_value ServerTickAsync(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecS *r = recs(env, a[0]);
  if (!r) return nullptr;
  aero3d_server *s = r->s;
  AsyncCall *c = new AsyncCall();
  auto counts = std::make_shared<std::pair<int, int>>(0, 0);
  c->run = [s, counts](std::string &m) {
    int rc = aero3d_server_tick(s, &counts->first, &counts->second);
    if (rc) m = "libaero3d batch error " + std::to_string(rc) + ": " + aero3d_batch_last_error(aero3d_server_batch(s));
    return rc;
  };
  c->result = [counts](napi_env e) { return tick_value(e, counts->first, counts->second); };
  return start_async(env, a[0], &r->batch.busy, c);
}
// serverPoll(h) -> [{tag, iters, converged, diverged, failed, nreads, res0, res, CL, CD, Cm, arrival, tAdmit, tDone}]
napi_value ServerPoll(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecS *r = recs(env, a[0]);
  if (!r) return nullptr;
  std::vector<aero3d_result> out(256);
  napi_value arr;
  NAPI_OK(napi_create_array(env, &arr));
  uint32_t k = 0;
  for (;;) {
    int n = 0;
    int rc = aero3d_server_poll(r->s, out.data(), (int)out.size(), &n);
    if (rc) return batch_throw(env, aero3d_server_batch(r->s), rc);
    for (int i = 0; i < n; i++) {
      const aero3d_result &x = out[i];
      napi_value o;
      NAPI_OK(napi_create_object(env, &o));
      set_num(env, o, "tag", (double)x.tag); set_num(env, o, "iters", x.iters); set_num(env, o, "converged", x.converged);
      set_num(env, o, "diverged", x.diverged); set_num(env, o, "failed", x.failed); set_num(env, o, "nreads", x.nreads);
      set_num(env, o, "res0", x.res0); set_num(env, o, "res", x.res); set_num(env, o, "CL", x.CL); set_num(env, o, "CD", x.CD);
      set_num(env, o, "Cm", x.Cm); set_num(env, o, "arrival", x.arrival); set_num(env, o, "tAdmit", x.t_admit); set_num(env, o, "tDone", x.t_done);
      NAPI_OK(napi_set_element(env, arr, k++, o));
    }
    if (n < (int)out.size()) break;
  }
  return arr;
}
napi_value ServerClock(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecS *r = recs(env, a[0]);
  if (!r) return nullptr;
  double t;
  int rc = aero3d_server_clock(r->s, &t);
  if (rc) return batch_throw(env, aero3d_server_batch(r->s), rc);
  napi_value v;
  NAPI_OK(napi_create_double(env, t, &v));
  return v;
}
napi_value ServerStats(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecS *r = recs(env, a[0]);
  if (!r) return nullptr;
  double busy, gpu; int64_t cu;
  int rc = aero3d_server_stats(r->s, &busy, &gpu, &cu);
  if (rc) return batch_throw(env, aero3d_server_batch(r->s), rc);
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  set_num(env, o, "busyS", busy); set_num(env, o, "gpuS", gpu); set_num(env, o, "cellUpdates", (double)cu);
  return o;
}

// ---------------------------------------------------------------------------
// axisymmetric solver (axi.h)
// ---------------------------------------------------------------------------
napi_value axi_throw(napi_env env, axi_solver *s, int rc) {
  std::string msg = "libaero3d axi error " + std::to_string(rc) + ": " + axi_last_error(s);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}
RecAxi *recaxi(napi_env env, napi_value v) { return record<RecAxi>(env, v, g_recaxi, "an axisymmetric solver handle"); }
axi_solver *axi_handle(napi_env env, napi_value v) {
  RecAxi *r = recaxi(env, v);
  return r ? r->s : nullptr;
}
bool get_prop(napi_env env, napi_value o, const char *k, napi_value *v) {
  bool has = false;
  if (napi_has_named_property(env, o, k, &has) != napi_ok || !has || napi_get_named_property(env, o, k, v) != napi_ok) {
    std::string m = std::string("aero3d addon: zones.") + k + " missing";
    napi_throw_type_error(env, nullptr, m.c_str());
    return false;
  }
  return true;
}
// zones object -> axi_zones (solid/comp Uint8Array, the rest Float32Array)
bool zones_of(napi_env env, napi_value o, size_t N, axi_zones *z) {
  napi_value v;
  size_t n;
  uint8_t *u8;
  if (!get_prop(env, o, "solid", &v) || !typed(env, v, napi_uint8_array, &u8, &n)) return false;
  z->solid = u8;
  if (n != N) { napi_throw_range_error(env, nullptr, "zones: length"); return false; }
  if (!get_prop(env, o, "comp", &v) || !typed(env, v, napi_uint8_array, &u8, &n)) return false;
  z->comp = u8;
  const char *fk[9] = {"phi", "ae", "an", "wgt", "lmix", "tgt", "rate", "yrate", "mfd"};
  const float **fp[9] = {&z->phi, &z->ae, &z->an, &z->wgt, &z->lmix, &z->tgt, &z->rate, &z->yrate, &z->mfd};
  for (int i = 0; i < 9; i++) {
    float *f;
    if (!get_prop(env, o, fk[i], &v) || !typed(env, v, napi_float32_array, &f, &n)) return false;
    if (n != N) { napi_throw_range_error(env, nullptr, "zones: length"); return false; }
    *fp[i] = f;
  }
  return true;
}
// axiCreate(xf, rf, zones, comps Float64Array(32), fs Float64Array(3), Q0 Float64Array(5N))
napi_value AxiCreate(napi_env env, napi_callback_info info) {
  napi_value a[6];
  if (!args(env, info, 6, a)) return nullptr;
  double *xf, *rf, *comps, *fs, *Q0;
  size_t nx, nr, nc, nf, nq;
  if (!typed(env, a[0], napi_float64_array, &xf, &nx) || !typed(env, a[1], napi_float64_array, &rf, &nr) ||
      !typed(env, a[3], napi_float64_array, &comps, &nc) || !typed(env, a[4], napi_float64_array, &fs, &nf) ||
      !typed(env, a[5], napi_float64_array, &Q0, &nq))
    return nullptr;
  if (nx < 3 || nr < 3 || nx > (1u << 20) || nr > (1u << 20)) { napi_throw_range_error(env, nullptr, "axiCreate: sizes"); return nullptr; }
  const size_t N = (nx - 1) * (nr - 1);
  if (nc != 32 || nf != 3 || nq != 5 * N) { napi_throw_range_error(env, nullptr, "axiCreate: sizes"); return nullptr; }
  axi_zones z;
  if (!zones_of(env, a[2], N, &z)) return nullptr;
  axi_grid g{(int)nx - 1, (int)nr - 1, xf, rf};
  axi_solver *s = nullptr;
  int rc = axi_create(&g, &z, comps, fs, Q0, &s);
  if (rc) return axi_throw(env, nullptr, rc);
  RecAxi *r = new RecAxi{s, N, (int)nx - 1, (int)nr - 1};
  { std::lock_guard<std::mutex> l(g_mu); g_recaxi.insert(r); }
  napi_value ext;
  if (napi_create_external(env, r, finaxi, nullptr, &ext) != napi_ok) {
    finaxi(env, r, nullptr);
    napi_throw_error(env, nullptr, "aero3d addon: Node-API call failed");
    return nullptr;
  }
  return ext;
}
napi_value AxiDispose(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecAxi *r = recaxi(env, a[0]);
  if (!r) return nullptr;
  axi_destroy(r->s);
  r->s = nullptr;                // the record lives until the external is collected
  return undef(env);
}
// axiSetOpt(h, cfl, order, visc, hf, global, dtRatio)
napi_value AxiSetOpt(napi_env env, napi_callback_info info) {
  napi_value a[7];
  if (!args(env, info, 7, a)) return nullptr;
  axi_solver *s = axi_handle(env, a[0]);
  if (!s) return nullptr;
  int rc = axi_set_opt(s, num(env, a[1]), (int)num(env, a[2]), (int)num(env, a[3]), num(env, a[4]), (int)num(env, a[5]), num(env, a[6]));
  return rc ? axi_throw(env, s, rc) : undef(env);
}
napi_value AxiSetComps(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  axi_solver *s = axi_handle(env, a[0]);
  double *c; size_t n;
  if (!s || !typed(env, a[1], napi_float64_array, &c, &n)) return nullptr;
  if (n != 32) { napi_throw_range_error(env, nullptr, "setComps: 32 values"); return nullptr; }
  int rc = axi_set_comps(s, c);
  return rc ? axi_throw(env, s, rc) : undef(env);
}
napi_value AxiSetZones(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!args(env, info, 3, a)) return nullptr;
  RecAxi *r = recaxi(env, a[0]);
  if (!r) return nullptr;
  axi_solver *s = r->s;
  if ((size_t)num(env, a[2]) != r->N) { napi_throw_range_error(env, nullptr, "setZones: cell count differs from the solver's"); return nullptr; }
  axi_zones z;
  if (!zones_of(env, a[1], r->N, &z)) return nullptr;
  int rc = axi_set_zones(s, &z);
  return rc ? axi_throw(env, s, rc) : undef(env);
}
napi_value AxiSetFs(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  axi_solver *s = axi_handle(env, a[0]);
  double *f; size_t n;
  if (!s || !typed(env, a[1], napi_float64_array, &f, &n)) return nullptr;
  if (n != 3) { napi_throw_range_error(env, nullptr, "setFs: r, u, p"); return nullptr; }
  int rc = axi_set_fs(s, f);
  return rc ? axi_throw(env, s, rc) : undef(env);
}
napi_value AxiSetState(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  RecAxi *r = recaxi(env, a[0]);
  double *q; size_t n;
  if (!r || !typed(env, a[1], napi_float64_array, &q, &n) || !check_len(env, n, 5 * r->N, "setState")) return nullptr;
  axi_solver *s = r->s;
  int rc = axi_set_state(s, q);
  return rc ? axi_throw(env, s, rc) : undef(env);
}
napi_value AxiStep(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  axi_solver *s = axi_handle(env, a[0]);
  if (!s) return nullptr;
  int rc = axi_step(s, (int)num(env, a[1]));
  return rc ? axi_throw(env, s, rc) : undef(env);
}
// axiRead(h, Q Float32Array(5N), src, mfx)
napi_value AxiRead(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!args(env, info, 4, a)) return nullptr;
  RecAxi *r = recaxi(env, a[0]);
  float *q, *sr, *mf;
  size_t nq, ns, nm;
  if (!r || !typed(env, a[1], napi_float32_array, &q, &nq) || !typed(env, a[2], napi_float32_array, &sr, &ns) ||
      !typed(env, a[3], napi_float32_array, &mf, &nm) || !check_len(env, nq, 5 * r->N, "axiRead Q") ||
      !check_len(env, ns, r->N, "axiRead src") || !check_len(env, nm, r->N, "axiRead mfx"))
    return nullptr;
  axi_solver *s = r->s;
  int rc = axi_read(s, q, sr, mf, nullptr);
  return rc ? axi_throw(env, s, rc) : undef(env);
}
// axiMonitorSections(h, cols Int32Array(n), rows Uint8Array(n H))
napi_value AxiMonitorSections(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!args(env, info, 3, a)) return nullptr;
  RecAxi *r = recaxi(env, a[0]);
  int32_t *cols; uint8_t *rows; size_t n, m;
  if (!r || !typed(env, a[1], napi_int32_array, &cols, &n) || !typed(env, a[2], napi_uint8_array, &rows, &m) ||
      !check_len(env, m, n * (size_t)r->H, "section rows"))
    return nullptr;
  int rc = axi_monitor_sections(r->s, (int)n, cols, rows);
  return rc ? axi_throw(env, r->s, rc) : undef(env);
}
// axiSections(h, cols Int32Array(n), rows Uint8Array(n H), sums Float64Array(11 n))
napi_value AxiSections(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!args(env, info, 4, a)) return nullptr;
  RecAxi *r = recaxi(env, a[0]);
  int32_t *cols; uint8_t *rows; double *o; size_t n, m, no;
  if (!r || !typed(env, a[1], napi_int32_array, &cols, &n) || !typed(env, a[2], napi_uint8_array, &rows, &m) ||
      !check_len(env, m, n * (size_t)r->H, "section rows") || !typed(env, a[3], napi_float64_array, &o, &no) ||
      !check_len(env, no, 11 * n, "section sums"))
    return nullptr;
  int rc = axi_sections(r->s, (int)n, cols, rows, o);
  return rc ? axi_throw(env, r->s, rc) : undef(env);
}
// axiSetDisplayZones(h, stream Int8Array(N), mat Uint8Array(N))
napi_value AxiSetDisplayZones(napi_env env, napi_callback_info info) {
  napi_value a[3];
  if (!args(env, info, 3, a)) return nullptr;
  RecAxi *r = recaxi(env, a[0]);
  int8_t *st; uint8_t *mat; size_t n, m;
  if (!r || !typed(env, a[1], napi_int8_array, &st, &n) || !check_len(env, n, r->N, "stream") ||
      !typed(env, a[2], napi_uint8_array, &mat, &m) || !check_len(env, m, r->N, "mat"))
    return nullptr;
  int rc = axi_set_display_zones(r->s, st, mat);
  return rc ? axi_throw(env, r->s, rc) : undef(env);
}
// axiDerived(h, disp, T0, P0, fields Float32Array(13 N) | null, reg Uint8Array(N) | null, mat Uint8Array(N) | null)
napi_value AxiDerived(napi_env env, napi_callback_info info) {
  napi_value a[7];
  if (!args(env, info, 7, a)) return nullptr;
  RecAxi *r = recaxi(env, a[0]);
  if (!r) return nullptr;
  float *f; uint8_t *reg, *mat;
  if (!opt_typed(env, a[4], napi_float32_array, 13 * r->N, &f, "display fields") ||
      !opt_typed(env, a[5], napi_uint8_array, r->N, &reg, "reg") || !opt_typed(env, a[6], napi_uint8_array, r->N, &mat, "mat"))
    return nullptr;
  int rc = axi_derived(r->s, (int)num(env, a[1]), num(env, a[2]), num(env, a[3]), f, reg, mat);
  return rc ? axi_throw(env, r->s, rc) : undef(env);
}
// axiTransfer(hdst, hsrc, T0, P0)
napi_value AxiTransfer(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!args(env, info, 4, a)) return nullptr;
  RecAxi *d = recaxi(env, a[0]);
  if (!d) return nullptr;
  RecAxi *s = recaxi(env, a[1]);
  if (!s) return nullptr;
  int rc = axi_transfer(d->s, s->s, num(env, a[2]), num(env, a[3]));
  return rc ? axi_throw(env, d->s, rc) : undef(env);
}
// axiMonitor(h, comps Float64Array(32), P0, out Float64Array(29)):
// F Fb Fw mfuel pw[8] res_sum2 res_n res_valid nsec (mdot w wTt) x 4 bad
napi_value AxiMonitor(napi_env env, napi_callback_info info) {
  napi_value a[4];
  if (!args(env, info, 4, a)) return nullptr;
  RecAxi *r = recaxi(env, a[0]);
  double *c, *o; size_t nc, no;
  if (!r || !typed(env, a[1], napi_float64_array, &c, &nc) || !check_len(env, nc, 32, "comps") ||
      !typed(env, a[3], napi_float64_array, &o, &no) || !check_len(env, no, 29, "monitor output"))
    return nullptr;
  axi_monitor_out m;
  int rc = axi_monitor(r->s, c, num(env, a[2]), &m);
  if (rc) return axi_throw(env, r->s, rc);
  o[0] = m.F; o[1] = m.Fb; o[2] = m.Fw; o[3] = m.mfuel;
  for (int q = 0; q < 8; q++) o[4 + q] = m.pw[q];
  o[12] = m.res_sum2; o[13] = m.res_n; o[14] = m.res_valid; o[15] = m.nsec;
  for (int q = 0; q < 4; q++) { o[16 + 3 * q] = m.sec_mdot[q]; o[17 + 3 * q] = m.sec_w[q]; o[18 + 3 * q] = m.sec_wTt[q]; }
  o[28] = m.bad;
  return undef(env);
}
// axiSetFused(h, form): kernel form of the uncapped steps (0 face + cell, 1 .. 4 fused tiles)
napi_value AxiSetFused(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  axi_solver *s = axi_handle(env, a[0]);
  if (!s) return nullptr;
  int rc = axi_set_fused(s, (int)num(env, a[1]));
  return rc ? axi_throw(env, s, rc) : undef(env);
}
// axiAutotune(h, iters) -> {form, msPerIter}
napi_value AxiAutotune(napi_env env, napi_callback_info info) {
  napi_value a[2];
  if (!args(env, info, 2, a)) return nullptr;
  axi_solver *s = axi_handle(env, a[0]);
  if (!s) return nullptr;
  int form = 0; float ms = 0;
  int rc = axi_autotune(s, (int)num(env, a[1]), &form, &ms);
  if (rc) return axi_throw(env, s, rc);
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  set_num(env, o, "form", form); set_num(env, o, "msPerIter", ms);
  return o;
}
napi_value AxiSync(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  axi_solver *s = axi_handle(env, a[0]);
  if (!s) return nullptr;
  int rc = axi_sync(s);
  return rc ? axi_throw(env, s, rc) : undef(env);
}

// ---- engine field model (axi_field_*, axi_initial_state)
RecField *recfield(napi_env env, napi_value v) { return record<RecField>(env, v, g_recfield, "a field model handle"); }
napi_value field_throw(napi_env env, axi_field *f, int rc) {
  std::string msg = "libaero3d field error " + std::to_string(rc) + ": " + axi_field_last_error(f);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}
// fieldCreate(table Float64Array) -> external handle
napi_value FieldCreate(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  double *t; size_t n;
  if (!typed(env, a[0], napi_float64_array, &t, &n)) return nullptr;
  if (n > (1u << 26)) { napi_throw_range_error(env, nullptr, "fieldCreate: table too long"); return nullptr; }
  axi_field *f = nullptr;
  int rc = axi_field_create(t, (int)n, &f);
  if (rc) return field_throw(env, nullptr, rc);
  RecField *r = new RecField{f};
  { std::lock_guard<std::mutex> l(g_mu); g_recfield.insert(r); }
  napi_value ext;
  if (napi_create_external(env, r, finfield, nullptr, &ext) != napi_ok) {
    finfield(env, r, nullptr);
    napi_throw_error(env, nullptr, "aero3d addon: Node-API call failed");
    return nullptr;
  }
  return ext;
}
napi_value FieldDispose(napi_env env, napi_callback_info info) {
  napi_value a[1];
  if (!args(env, info, 1, a)) return nullptr;
  RecField *r = recfield(env, a[0]);
  if (!r) return nullptr;
  axi_field_destroy(r->s);
  r->s = nullptr;
  return undef(env);
}
// fieldSample(h, x, r, th, out Float64Array(14 n))
napi_value FieldSample(napi_env env, napi_callback_info info) {
  napi_value a[5];
  if (!args(env, info, 5, a)) return nullptr;
  RecField *r = recfield(env, a[0]);
  double *x, *rr, *th, *o; size_t nx, nr, nt, no;
  if (!r || !typed(env, a[1], napi_float64_array, &x, &nx) || !typed(env, a[2], napi_float64_array, &rr, &nr) ||
      !typed(env, a[3], napi_float64_array, &th, &nt) || !typed(env, a[4], napi_float64_array, &o, &no) ||
      !check_len(env, nr, nx, "fieldSample r") || !check_len(env, nt, nx, "fieldSample th") || !check_len(env, no, 14 * nx, "fieldSample out"))
    return nullptr;
  if (nx > (1u << 26)) { napi_throw_range_error(env, nullptr, "fieldSample: at most 2^26 points"); return nullptr; }
  int rc = axi_field_sample(r->s, x, rr, th, (int)nx, o);
  return rc ? field_throw(env, r->s, rc) : undef(env);
}
// initialState(h, xf, rf, solid Uint8Array(N), stream Int8Array(N), uniform, Q Float64Array(5 N))
napi_value InitialState(napi_env env, napi_callback_info info) {
  napi_value a[7];
  if (!args(env, info, 7, a)) return nullptr;
  RecField *r = recfield(env, a[0]);
  double *xf, *rf, *Q; uint8_t *solid; int8_t *stream; size_t nx, nr, ns, nt, nq;
  if (!r || !typed(env, a[1], napi_float64_array, &xf, &nx) || !typed(env, a[2], napi_float64_array, &rf, &nr) ||
      !typed(env, a[3], napi_uint8_array, &solid, &ns) || !typed(env, a[4], napi_int8_array, &stream, &nt) ||
      !typed(env, a[6], napi_float64_array, &Q, &nq))
    return nullptr;
  if (nx < 2 || nr < 2 || nx > (1u << 20) || nr > (1u << 20)) { napi_throw_range_error(env, nullptr, "initialState: sizes"); return nullptr; }
  const size_t N = (nx - 1) * (nr - 1);
  if (!check_len(env, ns, N, "initialState solid") || !check_len(env, nt, N, "initialState stream") || !check_len(env, nq, 5 * N, "initialState Q"))
    return nullptr;
  bool uni = false;
  napi_get_value_bool(env, a[5], &uni);
  int rc = axi_initial_state(r->s, (int)nx - 1, (int)nr - 1, xf, rf, solid, stream, uni ? 1 : 0, Q);
  return rc ? field_throw(env, r->s, rc) : undef(env);
}

napi_value Init(napi_env env, napi_value exports) {
  struct { const char *name; napi_callback cb; } fns[] = {
      {"create", Create}, {"dispose", Dispose}, {"setFree", SetFree}, {"setFix", SetFix}, {"setOpt", SetOpt},
      {"reset", Reset}, {"step", Step}, {"iter", Iter}, {"read", Read}, {"write", Write}, {"sync", Sync},
      {"setForceList", SetForceList}, {"forces", Forces}, {"trefftz", Trefftz}, {"residual", Residual},
      {"timeSteps", TimeSteps}, {"autotune", Autotune}, {"readback", Readback}, {"probe", Probe}, {"stepAsync", StepAsync},
      {"batchCreate", BatchCreate}, {"batchDispose", BatchDispose}, {"batchAddGeometry", BatchAddGeometry},
      {"batchRemoveGeometry", BatchRemoveGeometry}, {"batchAdd", BatchAdd}, {"batchRemove", BatchRemove}, {"batchStep", BatchStep},
      {"batchStepAsync", BatchStepAsync}, {"batchSync", BatchSync}, {"batchIter", BatchIter}, {"batchRead", BatchRead},
      {"batchResiduals", BatchResiduals}, {"batchForces", BatchForces}, {"batchSetOption", BatchSetOption},
      {"batchFootprint", BatchFootprint},
      {"serverCreate", ServerCreate}, {"serverDispose", ServerDispose}, {"serverAddGeometry", ServerAddGeometry},
      {"serverRemoveGeometry", ServerRemoveGeometry}, {"serverSubmit", ServerSubmit}, {"serverTick", ServerTick},
      {"serverTickAsync", ServerTickAsync}, {"serverPoll", ServerPoll}, {"serverClock", ServerClock}, {"serverStats", ServerStats},
      {"axiCreate", AxiCreate}, {"axiDispose", AxiDispose}, {"axiSetOpt", AxiSetOpt}, {"axiSetComps", AxiSetComps},
      {"axiSetZones", AxiSetZones}, {"axiSetFs", AxiSetFs}, {"axiSetState", AxiSetState}, {"axiStep", AxiStep},
      {"axiRead", AxiRead}, {"axiSync", AxiSync}, {"axiSetFused", AxiSetFused}, {"axiAutotune", AxiAutotune}, {"axiMonitorSections", AxiMonitorSections}, {"axiMonitor", AxiMonitor},
      {"axiSections", AxiSections}, {"axiSetDisplayZones", AxiSetDisplayZones}, {"axiDerived", AxiDerived}, {"axiTransfer", AxiTransfer},
      {"fieldCreate", FieldCreate}, {"fieldDispose", FieldDispose}, {"fieldSample", FieldSample}, {"initialState", InitialState}};
  for (auto &f : fns) {
    napi_value fn;
    NAPI_OK(napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn));
    NAPI_OK(napi_set_named_property(env, exports, f.name, fn));
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
