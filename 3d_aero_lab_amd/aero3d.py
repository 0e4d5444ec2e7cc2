"""Python binding of libaero3d.so (include/aero3d.h) for tests and benchmarks.

`CfdHIP` follows the reference solver-object protocol of CfdCPU / CfdGPU
(src/air_cfd.js:157-259, 367-438): set_free, set_fix, opt (cfl, order), reset,
step(n), read(), dispose(); plus the device reductions forces(), trefftz() and
residual(). The HIP library is the only implementation: loading fails loudly
when it is missing or no device is usable.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# AERO3D_LIB names another build of the library (default: the in-tree one)
LIB_PATH = os.environ.get("AERO3D_LIB") or os.path.join(_HERE, "lib", "libaero3d.so")

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

EXPORTS = ["aero3d_create", "aero3d_destroy", "aero3d_last_error", "aero3d_set_free", "aero3d_set_fix",
           "aero3d_set_opt", "aero3d_reset", "aero3d_step", "aero3d_iter", "aero3d_read", "aero3d_write",
           "aero3d_sync", "aero3d_set_force_list", "aero3d_forces", "aero3d_trefftz", "aero3d_residual",
           "aero3d_time_steps", "aero3d_set_tile", "aero3d_get_tile", "aero3d_autotune", "aero3d_footprint",
           "aero3d_readback", "aero3d_readback_static", "aero3d_probe"]

PROBE_STATE_KEYS = ["P", "T", "rho", "V", "M", "Tt", "Pt", "Cp", "Taw", "Vx", "Vy", "Vz"]


class Grid(C.Structure):
    _fields_ = [("NX", C.c_int), ("NY", C.c_int), ("NZ", C.c_int),
                ("fx", C.POINTER(C.c_double)), ("fy", C.POINTER(C.c_double)), ("fz", C.POINTER(C.c_double))]


class ReadbackIn(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("alpha", "M", "S", "xcg", "zcg", "cref", "xT")]


class ReadbackOut(C.Structure):
    _fields_ = [("it", C.c_int)] + [(k, C.c_double) for k in ("res", "CL", "CD", "Cm", "CDi", "Mmax")]


class ProbeFs(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("T", "a", "P", "rho", "q")]


class ProbeOut(C.Structure):
    _fields_ = [("idx", C.c_void_p), ("w", C.c_void_p), ("ok", C.c_void_p), ("W", C.c_void_p), ("state", C.c_void_p),
                ("vel", C.c_void_p), ("vok", C.c_void_p)]


class Field(C.Structure):
    """aero3d_field: a state on the device in arrays the caller owns (device
    addresses as integers, xc on the host)."""
    _fields_ = [("NX", C.c_int), ("NY", C.c_int), ("NZ", C.c_int), ("ns", C.c_uint64), ("Q", C.c_void_p), ("rsnap", C.c_void_p),
                ("nit", C.c_int), ("t", C.c_void_p), ("dy", C.c_void_p), ("dz", C.c_void_p), ("xc", C.c_void_p),
                ("fq", C.c_void_p), ("cx", C.c_void_p), ("cz", C.c_void_p), ("x", C.c_void_p), ("z", C.c_void_p), ("nf", C.c_int)]


def readback_static(field, alpha, M, S, xcg, zcg, cref, xT=float("nan"), stream=None):
    """aero3d_readback_static: the readback over caller-owned device arrays, with no allocation."""
    o = ReadbackOut()
    rc = lib().aero3d_readback_static(C.byref(field), C.byref(ReadbackIn(alpha, M, S, xcg, zcg, cref, xT)), C.byref(o), stream)
    if rc:
        raise Aero3DError(f"aero3d_readback_static failed ({rc}): {lib().aero3d_last_error(None).decode()}")
    return {k: getattr(o, k) for k, _ in ReadbackOut._fields_}


RULE_RES_DROP, RULE_REFERENCE = 0, 1


class ConvRule:
    """Convergence rules of a monitored solve, one update per readback: the
    host mirror of the device rules of the served form.

    RULE_RES_DROP: converged at res / res0 <= res_drop (res0: the first
    positive residual), done then or at `cap` iterations. RULE_REFERENCE: the
    reference controller's rule (CFD3.readback, src/air_cfd.js:573-576): more
    than six readbacks, res / res0 < 2e-3 and the lift coefficients of the last
    four readbacks within 2e-3; done then or past `cap` iterations. Both end a
    solve whose residual is not finite after its first readback (diverged)."""

    def __init__(self, rule=RULE_REFERENCE, res_drop=1e-4, cap=30000):
        self.rule, self.res_drop, self.cap = rule, res_drop, cap
        self.res0 = float("nan")
        self.cl = []
        self.nreads = 0
        self.conv = self.diverged = self.done = False

    def update(self, res, CL, it):
        import math
        self.nreads += 1
        if not (self.res0 > 0) and res > 0:
            self.res0 = res
        self.cl = (self.cl + [CL])[-4:]
        self.diverged = self.nreads > 1 and not math.isfinite(res)
        if self.rule == RULE_REFERENCE:
            ok = self.nreads > 6 and self.res0 > 0 and res / self.res0 < 2e-3
            if ok:
                ok = not any(c != c for c in self.cl) and max(self.cl) - min(self.cl) < 2e-3
            self.conv = ok
            self.done = self.conv or self.diverged or it > self.cap
        else:
            self.conv = self.res0 > 0 and res / self.res0 <= self.res_drop
            self.done = self.conv or self.diverged or it >= self.cap
        return self.done


_lib = None


def lib():
    """Load libaero3d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python 3d_aero_lab_amd/build.py` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.aero3d_create.argtypes = [C.POINTER(Grid), u8p, f32p, f32p, f32p, f32p, f32p, C.POINTER(vp)]
        L.aero3d_destroy.argtypes = [vp]
        L.aero3d_last_error.argtypes = [vp]
        L.aero3d_last_error.restype = C.c_char_p
        L.aero3d_set_free.argtypes = [vp, f64p]
        L.aero3d_set_fix.argtypes = [vp, f64p]
        L.aero3d_set_opt.argtypes = [vp, C.c_double, C.c_int]
        L.aero3d_reset.argtypes = [vp]
        L.aero3d_step.argtypes = [vp, C.c_int]
        L.aero3d_iter.argtypes = [vp, C.POINTER(C.c_int)]
        L.aero3d_read.argtypes = [vp] + [C.c_void_p] * 6
        L.aero3d_write.argtypes = [vp] + [f32p] * 5
        L.aero3d_sync.argtypes = [vp]
        L.aero3d_set_force_list.argtypes = [vp, i32p, f64p, f64p, f64p, f64p, C.c_int]
        L.aero3d_forces.argtypes = [vp] + [C.c_double] * 6 + [f64p]
        L.aero3d_trefftz.argtypes = [vp] + [C.c_double] * 4 + [C.POINTER(C.c_double)]
        L.aero3d_residual.argtypes = [vp, C.POINTER(C.c_double)]
        L.aero3d_time_steps.argtypes = [vp, C.c_int, C.POINTER(C.c_float)]
        L.aero3d_footprint.argtypes = [vp, C.POINTER(C.c_int64)]
        L.aero3d_set_tile.argtypes = [vp, C.c_int, C.c_int]
        L.aero3d_get_tile.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.aero3d_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.aero3d_readback.argtypes = [vp, C.POINTER(ReadbackIn), C.POINTER(ReadbackOut)]
        L.aero3d_readback_static.argtypes = [C.POINTER(Field), C.POINTER(ReadbackIn), C.POINTER(ReadbackOut), vp]
        L.aero3d_probe.argtypes = [vp, C.c_int, f64p, C.POINTER(ProbeFs), C.c_double, C.POINTER(ProbeOut)]
        _lib = L
    return _lib


class Aero3DError(RuntimeError):
    pass


class CfdHIP:
    """Solver object on one GPU. grid: dict with fx, fy, fz (Float64 faces);
    geo: dict with t (u8), phi (f32), ap (4 f32 arrays)."""

    kind = "hip"

    def __init__(self, grid, geo):
        L = lib()
        self._fx = np.ascontiguousarray(grid["fx"], dtype=np.float64)
        self._fy = np.ascontiguousarray(grid["fy"], dtype=np.float64)
        self._fz = np.ascontiguousarray(grid["fz"], dtype=np.float64)
        self.NX, self.NY, self.NZ = len(self._fx) - 1, len(self._fy) - 1, len(self._fz) - 1
        self.N = self.NX * self.NY * self.NZ
        g = Grid(self.NX, self.NY, self.NZ, self._fx.ctypes.data_as(C.POINTER(C.c_double)),
                 self._fy.ctypes.data_as(C.POINTER(C.c_double)), self._fz.ctypes.data_as(C.POINTER(C.c_double)))
        a = lambda x, dt: np.ascontiguousarray(x, dtype=dt)
        t, phi = a(geo["t"], np.uint8), a(geo["phi"], np.float32)
        ap = [a(geo["ap"][d], np.float32) for d in range(4)]
        for x in [t, phi] + ap:
            if x.shape != (self.N,):
                raise ValueError(f"mask has {x.shape}, grid needs ({self.N},)")
        h = C.c_void_p()
        rc = L.aero3d_create(C.byref(g), t, phi, *ap, C.byref(h))
        if rc != 0:
            raise Aero3DError(f"aero3d_create failed ({rc}): {L.aero3d_last_error(None).decode()}")
        self.h = h
        self.opt = {"cfl": 0.7, "order": 2}
        self._opt_set = None

    def _chk(self, rc):
        if rc != 0:
            raise Aero3DError(f"libaero3d error {rc}: {lib().aero3d_last_error(self.h).decode()}")

    def dispose(self):
        if getattr(self, "h", None):
            lib().aero3d_destroy(self.h)
            self.h = None

    __del__ = dispose

    def set_free(self, F):
        self._chk(lib().aero3d_set_free(self.h, np.asarray(F, dtype=np.float64)))

    def set_fix(self, fix):
        self._chk(lib().aero3d_set_fix(self.h, np.ascontiguousarray(np.asarray(fix, dtype=np.float64).reshape(15))))

    def _sync_opt(self):
        o = (float(self.opt["cfl"]), int(self.opt["order"]))
        if o != self._opt_set:
            self._chk(lib().aero3d_set_opt(self.h, *o))
            self._opt_set = o

    def reset(self):
        self._sync_opt()
        self._chk(lib().aero3d_reset(self.h))

    def step(self, n):
        self._sync_opt()
        self._chk(lib().aero3d_step(self.h, int(n)))

    def sync(self):
        self._chk(lib().aero3d_sync(self.h))

    @property
    def it(self):
        v = C.c_int()
        self._chk(lib().aero3d_iter(self.h, C.byref(v)))
        return v.value

    def read(self, with_dt=False):
        out = [np.empty(self.N, np.float32) for _ in range(6 if with_dt else 5)]
        ptrs = [o.ctypes.data_as(C.c_void_p) for o in out] + [None] * (6 - len(out))
        self._chk(lib().aero3d_read(self.h, *ptrs))
        return np.stack(out)

    def write(self, Q):
        self._chk(lib().aero3d_write(self.h, *[np.ascontiguousarray(Q[c], dtype=np.float32) for c in range(5)]))

    def set_force_list(self, fl):
        self._fl = {k: np.ascontiguousarray(fl[k], dtype=np.int32 if k == "q" else np.float64) for k in ("q", "cx", "cz", "x", "z")}
        f = self._fl
        self._chk(lib().aero3d_set_force_list(self.h, f["q"], f["cx"], f["cz"], f["x"], f["z"], len(f["q"])))

    def forces(self, alpha, M, S, xcg, zcg, cref):
        out = np.zeros(3, np.float64)
        self._chk(lib().aero3d_forces(self.h, alpha, M, S, xcg, zcg, cref, out))
        return {"CL": out[0], "CD": out[1], "Cm": out[2]}

    def trefftz(self, xT, alpha, M, S):
        v = C.c_double()
        self._chk(lib().aero3d_trefftz(self.h, xT, alpha, M, S, C.byref(v)))
        return v.value

    def residual(self):
        v = C.c_double()
        self._chk(lib().aero3d_residual(self.h, C.byref(v)))
        return v.value

    def readback(self, alpha, M, S, xcg, zcg, cref, xT=float("nan")):
        """CFD3.readback on the device: residual since the previous readback,
        coefficients, Trefftz CDi at xT (NaN: none), Mach maximum."""
        o = ReadbackOut()
        self._chk(lib().aero3d_readback(self.h, C.byref(ReadbackIn(alpha, M, S, xcg, zcg, cref, xT)), C.byref(o)))
        return {k: getattr(o, k) for k, _ in ReadbackOut._fields_}

    def probe(self, pts, fs=None, M=0.0):
        """Display sampling at points pts (n x 3, full model) on the device:
        the reference's cfdStencil / cfdSample / cfdState / cfdVelAt. fs: the
        freestream record {T, a, P, rho, q} of the dimensional state (None:
        no state). Returns idx (n x 8), w (n x 8), ok (n), W (n x 5), state
        (n x 12, columns PROBE_STATE_KEYS, or None), vel (n x 3), vok (n)."""
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = pts.shape[0]
        r = {"idx": np.zeros((n, 8), np.int32), "w": np.zeros((n, 8), np.float32), "ok": np.zeros(n, np.uint8),
             "W": np.zeros((n, 5), np.float64), "state": np.zeros((n, 12), np.float64) if fs is not None else None,
             "vel": np.zeros((n, 3), np.float64), "vok": np.zeros(n, np.uint8)}
        o = ProbeOut(*[r[k].ctypes.data if r[k] is not None else None for k in ("idx", "w", "ok", "W", "state", "vel", "vok")])
        f = C.byref(ProbeFs(*[fs[k] for k in ("T", "a", "P", "rho", "q")])) if fs is not None else None
        self._chk(lib().aero3d_probe(self.h, n, pts.reshape(-1), f, M, C.byref(o)))
        return r

    def solve(self, ref, rule, every=50, xT=float("nan")):
        """Step to convergence under `rule` (a ConvRule), one readback every
        `every` iterations (and at the cap of the residual-drop rule); the
        first takes the density snapshot only, as the reference controller's
        first readback does. Returns the readback history. ref: (alpha, M, S,
        xcg, zcg, cref)."""
        hist = []
        while not rule.done:
            it = self.it
            nxt = (it // every + 1) * every
            if rule.rule == RULE_RES_DROP:
                nxt = min(nxt, rule.cap)
            self.step(nxt - it)
            hist.append(self.readback(*ref, xT))
            rule.update(hist[-1]["res"], hist[-1]["CL"], hist[-1]["it"])
        return hist

    def time_steps(self, n):
        self._sync_opt()
        ms = C.c_float()
        self._chk(lib().aero3d_time_steps(self.h, int(n), C.byref(ms)))
        return ms.value

    def set_tile(self, ty, zchunk):
        self._chk(lib().aero3d_set_tile(self.h, ty, zchunk))

    def get_tile(self):
        a, b = C.c_int(), C.c_int()
        self._chk(lib().aero3d_get_tile(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, iters=10):
        """Pick the fastest stage-kernel shape for this grid; returns (ty, zchunk, ms_per_iter)."""
        self._sync_opt()
        a, b, ms = C.c_int(), C.c_int(), C.c_float()
        self._chk(lib().aero3d_autotune(self.h, iters, C.byref(a), C.byref(b), C.byref(ms)))
        return a.value, b.value, ms.value

    def footprint(self):
        v = C.c_int64()
        self._chk(lib().aero3d_footprint(self.h, C.byref(v)))
        return v.value


# ---------------------------------------------------------------------------
# batched and served forms
# ---------------------------------------------------------------------------
class Request(C.Structure):
    _fields_ = [("geometry", C.c_int), ("F", C.c_double * 5), ("fix", C.c_double * 15), ("cfl", C.c_double),
                ("order", C.c_int), ("alpha", C.c_double), ("M", C.c_double), ("S", C.c_double), ("xcg", C.c_double),
                ("zcg", C.c_double), ("cref", C.c_double), ("rule", C.c_int), ("res_drop", C.c_double), ("max_iters", C.c_int),
                ("arrival", C.c_double), ("tag", C.c_int64), ("state_out", C.c_void_p)]


class Result(C.Structure):
    _fields_ = [("tag", C.c_int64), ("iters", C.c_int), ("converged", C.c_int), ("diverged", C.c_int), ("failed", C.c_int),
                ("nreads", C.c_int), ("res0", C.c_double), ("res", C.c_double),
                ("CL", C.c_double), ("CD", C.c_double), ("Cm", C.c_double), ("arrival", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]


BATCH_EXPORTS = ["aero3d_batch_create", "aero3d_batch_destroy", "aero3d_batch_last_error", "aero3d_batch_add_geometry",
                 "aero3d_batch_remove_geometry", "aero3d_batch_case_bytes", "aero3d_batch_add", "aero3d_batch_remove",
                 "aero3d_batch_set_zchunk", "aero3d_batch_set_tile_width", "aero3d_batch_set_group_cells", "aero3d_batch_set_persistent", "aero3d_batch_step", "aero3d_batch_sync", "aero3d_batch_iter", "aero3d_batch_read",
                 "aero3d_batch_residuals", "aero3d_batch_forces", "aero3d_batch_footprint", "aero3d_batch_time_steps",
                 "aero3d_server_create", "aero3d_server_destroy", "aero3d_server_batch", "aero3d_server_clock",
                 "aero3d_server_submit", "aero3d_server_tick", "aero3d_server_poll", "aero3d_server_stats"]
EXPORTS = EXPORTS + BATCH_EXPORTS

_blib = None


def blib():
    global _blib
    if _blib is None:
        L = lib()
        vp, ci, cd, pi, pd = C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_double)
        p64 = C.POINTER(C.c_int64)
        L.aero3d_batch_create.argtypes = [C.c_int64, ci, C.POINTER(vp)]
        L.aero3d_batch_destroy.argtypes = [vp]
        L.aero3d_batch_last_error.argtypes = [vp]
        L.aero3d_batch_last_error.restype = C.c_char_p
        L.aero3d_batch_add_geometry.argtypes = [vp, C.POINTER(Grid), u8p, f32p, f32p, f32p, f32p, f32p, i32p, f64p, f64p, f64p, f64p, ci, pi]
        L.aero3d_batch_remove_geometry.argtypes = [vp, ci]
        L.aero3d_batch_case_bytes.argtypes = [vp, ci, p64]
        L.aero3d_batch_add.argtypes = [vp, ci, f64p, f64p, cd, ci, pi]
        L.aero3d_batch_remove.argtypes = [vp, ci]
        L.aero3d_batch_set_zchunk.argtypes = [vp, ci]
        L.aero3d_batch_set_tile_width.argtypes = [vp, ci]
        L.aero3d_batch_set_group_cells.argtypes = [vp, C.c_int64]
        L.aero3d_batch_set_persistent.argtypes = [vp, C.c_int]
        L.aero3d_batch_step.argtypes = [vp, ci]
        L.aero3d_batch_sync.argtypes = [vp]
        L.aero3d_batch_iter.argtypes = [vp, ci, pi]
        L.aero3d_batch_read.argtypes = [vp, ci] + [f32p] * 5
        L.aero3d_batch_residuals.argtypes = [vp, i32p, ci, f64p]
        L.aero3d_batch_forces.argtypes = [vp, ci] + [cd] * 6 + [f64p]
        L.aero3d_batch_footprint.argtypes = [vp, p64, p64]
        L.aero3d_batch_time_steps.argtypes = [vp, ci, C.POINTER(C.c_float)]
        L.aero3d_server_create.argtypes = [C.c_int64, ci, ci, C.POINTER(vp)]
        L.aero3d_server_destroy.argtypes = [vp]
        L.aero3d_server_batch.argtypes = [vp]
        L.aero3d_server_batch.restype = vp
        L.aero3d_server_clock.argtypes = [vp, pd]
        L.aero3d_server_submit.argtypes = [vp, C.POINTER(Request)]
        L.aero3d_server_tick.argtypes = [vp, pi, pi]
        L.aero3d_server_poll.argtypes = [vp, C.POINTER(Result), ci, pi]
        L.aero3d_server_stats.argtypes = [vp, pd, pd, p64]
        _blib = L
    return _blib


def _grid_struct(grid):
    fx, fy, fz = (np.ascontiguousarray(grid[k], dtype=np.float64) for k in ("fx", "fy", "fz"))
    g = Grid(len(fx) - 1, len(fy) - 1, len(fz) - 1, fx.ctypes.data_as(C.POINTER(C.c_double)),
             fy.ctypes.data_as(C.POINTER(C.c_double)), fz.ctypes.data_as(C.POINTER(C.c_double)))
    return g, (fx, fy, fz)


class Batch:
    """Many solves stepped together (aero3d_batch_*)."""

    def __init__(self, arena_bytes=8 << 30, max_slots=256, handle=None):
        L = blib()
        if handle is not None:
            self.h, self._own = handle, False
        else:
            h = C.c_void_p()
            rc = L.aero3d_batch_create(int(arena_bytes), max_slots, C.byref(h))
            if rc:
                raise Aero3DError(f"aero3d_batch_create failed ({rc})")
            self.h, self._own = h, True
        self.N = {}

    def _chk(self, rc):
        if rc:
            raise Aero3DError(f"libaero3d batch error {rc}: {blib().aero3d_batch_last_error(self.h).decode()}")

    def dispose(self):
        if getattr(self, "h", None) and self._own:
            blib().aero3d_batch_destroy(self.h)
        self.h = None

    __del__ = dispose

    def add_geometry(self, grid, geo, fl=None):
        g, keep = _grid_struct(grid)
        a = lambda x, dt: np.ascontiguousarray(x, dtype=dt)
        if fl is None:
            fl = {"q": np.zeros(0, np.int32), **{k: np.zeros(0) for k in ("cx", "cz", "x", "z")}}
        gid = C.c_int()
        self._chk(blib().aero3d_batch_add_geometry(self.h, C.byref(g), a(geo["t"], np.uint8), a(geo["phi"], np.float32),
                                                   *[a(geo["ap"][d], np.float32) for d in range(4)], a(fl["q"], np.int32),
                                                   *[a(fl[k], np.float64) for k in ("cx", "cz", "x", "z")], len(fl["q"]), C.byref(gid)))
        self.N[gid.value] = g.NX * g.NY * g.NZ
        return gid.value

    def remove_geometry(self, gid):
        self._chk(blib().aero3d_batch_remove_geometry(self.h, gid))

    def case_bytes(self, gid):
        v = C.c_int64()
        self._chk(blib().aero3d_batch_case_bytes(self.h, gid, C.byref(v)))
        return v.value

    def add(self, gid, F, fix, cfl, order=2):
        s = C.c_int()
        self._chk(blib().aero3d_batch_add(self.h, gid, np.asarray(F, np.float64), np.ascontiguousarray(np.asarray(fix, np.float64).reshape(15)),
                                          cfl, order, C.byref(s)))
        self.N[("slot", s.value)] = self.N[gid]
        return s.value

    def remove(self, slot):
        self._chk(blib().aero3d_batch_remove(self.h, slot))

    def set_zchunk(self, zc):
        self._chk(blib().aero3d_batch_set_zchunk(self.h, zc))

    def set_tile_width(self, width):
        """64 or 32 columns for every case, 0 (default): per case by its x extent."""
        self._chk(blib().aero3d_batch_set_tile_width(self.h, width))

    def set_persistent(self, on):
        self._chk(blib().aero3d_batch_set_persistent(self.h, int(bool(on))))

    def set_group_cells(self, cells):
        self._chk(blib().aero3d_batch_set_group_cells(self.h, int(cells)))

    def step(self, n):
        self._chk(blib().aero3d_batch_step(self.h, n))

    def sync(self):
        self._chk(blib().aero3d_batch_sync(self.h))

    def iter(self, slot):
        v = C.c_int()
        self._chk(blib().aero3d_batch_iter(self.h, slot, C.byref(v)))
        return v.value

    def read(self, slot):
        N = self.N[("slot", slot)]
        out = [np.empty(N, np.float32) for _ in range(5)]
        self._chk(blib().aero3d_batch_read(self.h, slot, *out))
        return np.stack(out)

    def residuals(self, slots):
        sl = np.ascontiguousarray(slots, np.int32)
        out = np.zeros(len(sl), np.float64)
        self._chk(blib().aero3d_batch_residuals(self.h, sl, len(sl), out))
        return out

    def forces(self, slot, alpha, M, S, xcg, zcg, cref):
        out = np.zeros(3, np.float64)
        self._chk(blib().aero3d_batch_forces(self.h, slot, alpha, M, S, xcg, zcg, cref, out))
        return {"CL": out[0], "CD": out[1], "Cm": out[2]}

    def footprint(self):
        u, c = C.c_int64(), C.c_int64()
        self._chk(blib().aero3d_batch_footprint(self.h, C.byref(u), C.byref(c)))
        return u.value, c.value

    def time_steps(self, n):
        ms = C.c_float()
        self._chk(blib().aero3d_batch_time_steps(self.h, n, C.byref(ms)))
        return ms.value


class Server:
    """Continuous-batching server of solve requests (aero3d_server_*)."""

    def __init__(self, arena_bytes=8 << 30, max_slots=256, check_every=50):
        h = C.c_void_p()
        rc = blib().aero3d_server_create(int(arena_bytes), max_slots, check_every, C.byref(h))
        if rc:
            raise Aero3DError(f"aero3d_server_create failed ({rc})")
        self.h = h
        self.batch = Batch(handle=C.c_void_p(blib().aero3d_server_batch(h)))
        self._keep_out = {}

    def dispose(self):
        if getattr(self, "h", None):
            blib().aero3d_server_destroy(self.h)
            self.h = None

    __del__ = dispose

    def _chk(self, rc):
        if rc:
            raise Aero3DError(f"libaero3d server error {rc}: {blib().aero3d_batch_last_error(self.batch.h).decode()}")

    def clock(self):
        v = C.c_double()
        self._chk(blib().aero3d_server_clock(self.h, C.byref(v)))
        return v.value

    def submit(self, gid, F, fix, cfl, order, force_ref, res_drop, max_iters, arrival, tag, rule=RULE_RES_DROP, state_out=None):
        """state_out: optional float32 array of 5 x N that receives the final state."""
        r = Request()
        r.geometry = gid
        r.F[:] = list(F)
        r.fix[:] = list(np.asarray(fix, np.float64).reshape(15))
        r.cfl, r.order = cfl, order
        r.alpha, r.M, r.S, r.xcg, r.zcg, r.cref = force_ref
        r.rule, r.res_drop, r.max_iters, r.arrival, r.tag = rule, res_drop, max_iters, arrival, tag
        if state_out is not None:
            assert state_out.dtype == np.float32 and state_out.flags["C_CONTIGUOUS"] and state_out.size == 5 * self.batch.N[gid]
            self._keep_out[tag] = state_out
            r.state_out = state_out.ctypes.data
        self._chk(blib().aero3d_server_submit(self.h, C.byref(r)))

    def tick(self):
        a, p = C.c_int(), C.c_int()
        self._chk(blib().aero3d_server_tick(self.h, C.byref(a), C.byref(p)))
        return a.value, p.value

    def poll(self, max_n=1024):
        buf = (Result * max_n)()
        n = C.c_int()
        self._chk(blib().aero3d_server_poll(self.h, buf, max_n, C.byref(n)))
        out = [{f: getattr(buf[i], f) for f, _ in Result._fields_} for i in range(n.value)]
        for r in out:
            self._keep_out.pop(r["tag"], None)
        return out

    def stats(self):
        """(wall seconds with resident work, device seconds of the ticks, cell-updates of retired cases)"""
        b, g, c = C.c_double(), C.c_double(), C.c_int64()
        self._chk(blib().aero3d_server_stats(self.h, C.byref(b), C.byref(g), C.byref(c)))
        return b.value, g.value, c.value


# ---------------------------------------------------------------------------
# axisymmetric engine solver (include/axi.h)
# ---------------------------------------------------------------------------
class AxiGrid(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("xf", C.POINTER(C.c_double)), ("rf", C.POINTER(C.c_double))]


class AxiZones(C.Structure):
    _fields_ = [("solid", C.c_void_p), ("comp", C.c_void_p)] + [(k, C.c_void_p) for k in
                                                                 ("phi", "ae", "an", "wgt", "lmix", "tgt", "rate", "yrate", "mfd")]


AXI_EXPORTS = ["axi_create", "axi_destroy", "axi_last_error", "axi_set_opt", "axi_set_comps", "axi_set_zones", "axi_set_fs",
               "axi_set_state", "axi_step", "axi_iter", "axi_read", "axi_sync", "axi_time_steps", "axi_monitor",
               "axi_monitor_sections", "axi_set_fused", "axi_autotune", "axi_sections", "axi_set_display_zones", "axi_derived",
               "axi_transfer", "axi_field_create", "axi_field_destroy", "axi_field_last_error", "axi_field_sample",
               "axi_initial_state"]
AXI_DERIVED_KEYS = ["T", "Tt", "P", "Pt", "V", "M", "rho", "Vx", "Vth", "Vr", "Y", "vort", "ds"]


class AxiMonitorOut(C.Structure):
    _fields_ = [("F", C.c_double), ("Fb", C.c_double), ("Fw", C.c_double), ("mfuel", C.c_double), ("pw", C.c_double * 8),
                ("res_sum2", C.c_double), ("res_n", C.c_double), ("res_valid", C.c_int), ("nsec", C.c_int),
                ("sec_mdot", C.c_double * 4), ("sec_w", C.c_double * 4), ("sec_wTt", C.c_double * 4), ("bad", C.c_int)]
_alib = None


def alib():
    global _alib
    if _alib is None:
        L = lib()
        vp = C.c_void_p
        L.axi_create.argtypes = [C.POINTER(AxiGrid), C.POINTER(AxiZones), f64p, f64p, f64p, C.POINTER(vp)]
        L.axi_destroy.argtypes = [vp]
        L.axi_last_error.argtypes = [vp]
        L.axi_last_error.restype = C.c_char_p
        L.axi_set_opt.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double]
        L.axi_set_comps.argtypes = [vp, f64p]
        L.axi_set_zones.argtypes = [vp, C.POINTER(AxiZones)]
        L.axi_set_fs.argtypes = [vp, f64p]
        L.axi_set_state.argtypes = [vp, f64p]
        L.axi_step.argtypes = [vp, C.c_int]
        L.axi_iter.argtypes = [vp, C.POINTER(C.c_int)]
        L.axi_read.argtypes = [vp] + [C.c_void_p] * 4
        L.axi_sync.argtypes = [vp]
        L.axi_time_steps.argtypes = [vp, C.c_int, C.POINTER(C.c_float)]
        L.axi_monitor.argtypes = [vp, f64p, C.c_double, C.POINTER(AxiMonitorOut)]
        L.axi_monitor_sections.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p]
        L.axi_set_fused.argtypes = [vp, C.c_int]
        L.axi_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.axi_sections.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.axi_set_display_zones.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.axi_derived.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        L.axi_transfer.argtypes = [vp, vp, C.c_double, C.c_double]
        L.axi_field_create.argtypes = [f64p, C.c_int, C.POINTER(vp)]
        L.axi_field_destroy.argtypes = [vp]
        L.axi_field_last_error.argtypes = [vp]
        L.axi_field_last_error.restype = C.c_char_p
        L.axi_field_sample.argtypes = [vp, f64p, f64p, f64p, C.c_int, f64p]
        L.axi_initial_state.argtypes = [vp, C.c_int, C.c_int, f64p, f64p, u8p, C.c_void_p, C.c_int, f64p]
        _alib = L
    return _alib


class AxiHIP:
    """Axisymmetric solver object (GPUSolver protocol, src/cfd_gpu.js:207-323)."""

    ZF = ("phi", "ae", "an", "wgt", "lmix", "tgt", "rate", "yrate", "mfd")

    def __init__(self, xf, rf, zones, comps, fs, Q0, opt=None):
        L = alib()
        self._xf = np.ascontiguousarray(xf, np.float64)
        self._rf = np.ascontiguousarray(rf, np.float64)
        self.W, self.H = len(self._xf) - 1, len(self._rf) - 1
        self.N = self.W * self.H
        g = AxiGrid(self.W, self.H, self._xf.ctypes.data_as(C.POINTER(C.c_double)), self._rf.ctypes.data_as(C.POINTER(C.c_double)))
        z = self._zones(zones)
        Q = np.ascontiguousarray(np.asarray(Q0, np.float64).reshape(-1))
        if Q.size != 5 * self.N:
            raise ValueError("Q0 must hold 5 planes of W*H values")
        h = C.c_void_p()
        rc = L.axi_create(C.byref(g), C.byref(z), np.ascontiguousarray(comps, np.float64),
                          np.array([fs["r"], fs["u"], fs["p"]], np.float64), Q, C.byref(h))
        if rc:
            raise Aero3DError(f"axi_create failed ({rc}): {L.axi_last_error(None).decode()}")
        self.h = h
        self.opt = dict({"cfl": 0.6, "order": 2, "visc": True, "hf": 0.99 * 43e6, "global": False, "dtRatio": 0}, **(opt or {}))
        self._opt_set = None

    def _zones(self, zones):
        self._zkeep = {"solid": np.ascontiguousarray(zones["solid"], np.uint8), "comp": np.ascontiguousarray(zones["comp"], np.uint8)}
        for k in self.ZF:
            self._zkeep[k] = np.ascontiguousarray(zones[k], np.float32)
        for k, a in self._zkeep.items():
            if a.shape != (self.N,):
                raise ValueError(f"zones.{k} has {a.shape}")
        return AxiZones(*[self._zkeep[k].ctypes.data for k in ("solid", "comp") + self.ZF])

    def _chk(self, rc):
        if rc:
            raise Aero3DError(f"libaero3d axi error {rc}: {alib().axi_last_error(self.h).decode()}")

    def dispose(self):
        if getattr(self, "h", None):
            alib().axi_destroy(self.h)
            self.h = None

    __del__ = dispose

    def _sync_opt(self):
        o = self.opt
        key = (float(o["cfl"]), int(o["order"]), int(bool(o["visc"])), float(o["hf"] or 0), int(bool(o["global"])), float(o["dtRatio"] or 0))
        if key != self._opt_set:
            self._chk(alib().axi_set_opt(self.h, *key))
            self._opt_set = key

    def set_comps(self, c):
        self._chk(alib().axi_set_comps(self.h, np.ascontiguousarray(c, np.float64)))

    def set_fused(self, form):
        """0: face + cell kernels; 1 .. 4: one fused tile kernel per stage (tile shapes)."""
        self._chk(alib().axi_set_fused(self.h, int(form)))

    def autotune(self, iters=50):
        """Keep the fastest kernel form for this grid; returns (form, ms per iteration). The state is kept."""
        self._sync_opt()
        f, ms = C.c_int(), C.c_float()
        self._chk(alib().axi_autotune(self.h, iters, C.byref(f), C.byref(ms)))
        return f.value, ms.value

    def set_zones(self, zones):
        z = self._zones(zones)
        self._chk(alib().axi_set_zones(self.h, C.byref(z)))

    def set_fs(self, fs):
        self._chk(alib().axi_set_fs(self.h, np.array([fs["r"], fs["u"], fs["p"]], np.float64)))

    def set_state(self, Q):
        self._chk(alib().axi_set_state(self.h, np.ascontiguousarray(np.asarray(Q, np.float64).reshape(-1))))

    def step(self, n=1):
        self._sync_opt()
        self._chk(alib().axi_step(self.h, n))

    def read(self):
        Q = np.empty(5 * self.N, np.float32)
        src = np.empty(self.N, np.float32)
        mfx = np.empty(self.N, np.float32)
        dt = np.empty(self.N, np.float32)
        self._chk(alib().axi_read(self.h, Q.ctypes.data, src.ctypes.data, mfx.ctypes.data, dt.ctypes.data))
        return Q.reshape(5, self.N), src, mfx, dt

    @property
    def iter(self):
        v = C.c_int()
        self._chk(alib().axi_iter(self.h, C.byref(v)))
        return v.value

    def time_steps(self, n):
        self._sync_opt()
        ms = C.c_float()
        self._chk(alib().axi_time_steps(self.h, n, C.byref(ms)))
        return ms.value

    def monitor_sections(self, cols, rows):
        """cols: column per section; rows: (nsec, H) mask of the rows each section counts."""
        cols = np.ascontiguousarray(cols, np.int32)
        rows = np.ascontiguousarray(rows, np.uint8).reshape(-1)
        assert rows.size == cols.size * self.H
        self._chk(alib().axi_monitor_sections(self.h, int(cols.size), cols.ctypes.data, rows.ctypes.data))

    def monitor(self, comps, P0):
        o = AxiMonitorOut()
        self._chk(alib().axi_monitor(self.h, np.ascontiguousarray(comps, np.float64), P0, C.byref(o)))
        n = o.nsec
        return {"F": o.F, "Fb": o.Fb, "Fw": o.Fw, "mfuel": o.mfuel, "pw": list(o.pw), "res_sum2": o.res_sum2,
                "res_n": o.res_n, "res_valid": bool(o.res_valid), "bad": o.bad,
                "sec": [{"mdot": o.sec_mdot[q], "w": o.sec_w[q], "wTt": o.sec_wTt[q]} for q in range(n)]}

    def sections(self, cols, rows):
        """Section sums of the current state (axi_sections): (nsec, 11) of mass flow,
        w, area, sum p dA and the w-weighted Tt, Pt, T, p, |V|, M, Y."""
        cols = np.ascontiguousarray(cols, np.int32)
        rows = np.ascontiguousarray(rows, np.uint8).reshape(-1)
        assert rows.size == cols.size * self.H
        out = np.zeros(11 * cols.size)
        self._chk(alib().axi_sections(self.h, int(cols.size), cols.ctypes.data, rows.ctypes.data, out.ctypes.data))
        return out.reshape(-1, 11)

    def set_display_zones(self, stream, mat):
        st = np.ascontiguousarray(stream, np.int8)
        mt = np.ascontiguousarray(mat, np.uint8)
        assert st.shape == (self.N,) and mt.shape == (self.N,)
        self._chk(alib().axi_set_display_zones(self.h, st.ctypes.data, mt.ctypes.data))

    def derived(self, T0, P0, disp):
        """Display fields (axi_derived): (13 x N float32 in AXI_DERIVED_KEYS order, reg, mat)."""
        f = np.empty(13 * self.N, np.float32)
        reg = np.empty(self.N, np.uint8)
        mat = np.empty(self.N, np.uint8)
        self._chk(alib().axi_derived(self.h, int(disp), T0, P0, f.ctypes.data, reg.ctypes.data, mat.ctypes.data))
        return f.reshape(13, self.N), reg, mat

    def transfer_from(self, src, T0, P0):
        """Grid sequencing (axi_transfer): src's state carried onto this solver's grid."""
        self._chk(alib().axi_transfer(self.h, src.h, T0, P0))

# ---------------------------------------------------------------------------
# Engine field model (FieldModel, src/field.js:255-488) and initialState
# (src/cfd_grid.js:279-305) on the device.

FIELD_SAMPLE_KEYS = ["region", "T", "P", "Tt", "Pt", "V", "Vx", "Vth", "M", "a", "rho", "g", "cp", "R"]
_STREAM_NAME = {"inlet": 0, "bypass": 1, "core": 2, "ab": 3}
_ENGINE_TYPE = {"tj": 0, "hbtf": 1, "lbtf": 2}
FM_HEAD = 240


def field_table(d):
    """The flat table of axi_field_create (layout: AXI_FM_* in axi.h) from a field
    model description: the FieldModel constructor's products and the geometry it
    samples, as written by the model's owner (JSON: streams, kn, bodies, c, fl,
    comb, mixer, ab, lip, outer, outerEnd, cold, exits, Uface, maxdT, hotGas)."""
    T = [0.0] * FM_HEAD
    nan = float("nan")
    num = lambda v: nan if v is None else float(v)

    def put(vals):
        off = len(T)
        T.extend(float(v) for v in vals)
        return off

    def poly(pts):
        return put([c for p in pts for c in p]), len(pts)

    order = d["order"]
    T[0], T[1], T[2], T[3] = 1.0, _ENGINE_TYPE[d["type"]], d["s"], len(order)
    for i, k in enumerate(order):
        st = d["streams"][k]
        lo, hi = poly(st["lo"]), poly(st["hi"])
        kn = put([c for q in d["kn"][k] for c in q])
        T[4 + 10 * i: 14 + 10 * i] = [_STREAM_NAME[k], st["x0"], st["x1"], lo[0], lo[1], hi[0], hi[1], kn, len(d["kn"][k]), 0]
    T[44] = len(d["bodies"])
    for b, body in enumerate(d["bodies"]):
        xs, rs = [p[0] for p in body["poly"]], [p[1] for p in body["poly"]]
        off, n = poly(body["poly"])
        T[45 + 6 * b: 51 + 6 * b] = [min(xs), max(xs), min(rs), max(rs), off, n]
    cb = d["comb"]
    li, lo = poly(cb["linerIn"]), poly(cb["linerOut"])
    T[141:152] = [order.index(cb["stream"]), cb["x0"], cb["xd"], cb["x1"], cb["turbEnd"], cb["n"], 1.0 if cb["cannular"] else 0.0,
                  li[0], li[1], lo[0], lo[1]]
    mx = d.get("mixer")
    T[152:156] = [1.0, mx["x"], mx["coreFrac"], mx["lobes"]] if mx else [0.0] * 4
    ab = d.get("ab")
    if ab:
        rg = put(ab["rings"])
        T[156:162] = [1.0, order.index(ab["stream"]), ab["xFH"], ab["x1"], rg, len(ab["rings"])]
    c = d["c"]
    T[162:169] = [num(c["Tt3"]), num(c["Tt4"]), num(c["Tt5"]), num(c["Tt6A"]), num(c["Tt7"]), num(c["Tt13"]), 1.0 if c["abOn"] else 0.0]
    fl = d["fl"]
    T[169:174] = [fl["V0"], fl["T0"], fl["P0"], fl["Tt0"], fl["Pt0"]]
    ol = poly(d["outer"])
    T[174:180] = [d["lip"]["x"], d["lip"]["r"], d["lip"]["rOut"], ol[0], ol[1], d["outerEnd"]]
    T[180:182] = [d["cold"]["g"], d["cold"]["cp"]]
    T[182] = len(d["exits"])
    for q, e in enumerate(d["exits"]):
        T[183 + 12 * q: 195 + 12 * q] = [e["x"], e["R"], e["spread"], e["Qm"], e["Qh"], 1.0 if e["under"] else 0.0, e["Ls"],
                                         e["P"], e["Pt"], e["g"], e["cp"], e["dTt"]]
    hg = d.get("hotGas")
    T[231:236] = [d["Uface"], d["maxdT"], 1.0 if hg else 0.0, hg["g"] if hg else 0.0, hg["cp"] if hg else 0.0]
    return np.array(T, np.float64)


class FieldModelHIP:
    """The engine field model on the device (axi_field_*): sample() is
    FieldModel.prototype.sample at many points, initial_state() the solver's
    initialState. `desc` is a field model description (field_table)."""

    def __init__(self, desc):
        self.table = field_table(desc)
        self.order = list(desc["order"])
        h = C.c_void_p()
        rc = alib().axi_field_create(self.table, int(self.table.size), C.byref(h))
        if rc != 0:
            raise Aero3DError(rc, alib().axi_field_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc != 0:
            raise Aero3DError(rc, alib().axi_field_last_error(self.h).decode())

    def dispose(self):
        if getattr(self, "h", None):
            alib().axi_field_destroy(self.h)
            self.h = None

    __del__ = dispose

    def sample(self, x, r, th=None):
        """(n, 14) in FIELD_SAMPLE_KEYS order."""
        x = np.ascontiguousarray(x, np.float64)
        r = np.ascontiguousarray(r, np.float64)
        th = np.zeros_like(x) if th is None else np.ascontiguousarray(th, np.float64)
        assert x.shape == r.shape == th.shape and x.ndim == 1
        out = np.empty(14 * x.size)
        self._chk(alib().axi_field_sample(self.h, x, r, th, int(x.size), out))
        return out.reshape(-1, 14)

    def initial_state(self, xf, rf, solid, stream, uniform=False):
        """initialState on the grid of faces xf, rf: (5, W*H) float64."""
        xf = np.ascontiguousarray(xf, np.float64)
        rf = np.ascontiguousarray(rf, np.float64)
        W, H = xf.size - 1, rf.size - 1
        solid = np.ascontiguousarray(solid, np.uint8)
        stream = np.ascontiguousarray(stream, np.int8)
        assert solid.shape == (W * H,) and stream.shape == (W * H,)
        Q = np.empty(5 * W * H)
        self._chk(alib().axi_initial_state(self.h, W, H, xf, rf, solid, stream.ctypes.data, int(bool(uniform)), Q))
        return Q.reshape(5, W * H)


EXPORTS = EXPORTS + AXI_EXPORTS
