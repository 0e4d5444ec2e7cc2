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
LIB_PATH = os.path.join(_HERE, "lib", "libaero3d.so")

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

EXPORTS = ["aero3d_create", "aero3d_destroy", "aero3d_last_error", "aero3d_set_free", "aero3d_set_fix",
           "aero3d_set_opt", "aero3d_reset", "aero3d_step", "aero3d_iter", "aero3d_read", "aero3d_write",
           "aero3d_sync", "aero3d_set_force_list", "aero3d_forces", "aero3d_trefftz", "aero3d_residual",
           "aero3d_time_steps", "aero3d_set_tile", "aero3d_get_tile", "aero3d_autotune", "aero3d_footprint"]


class Grid(C.Structure):
    _fields_ = [("NX", C.c_int), ("NY", C.c_int), ("NZ", C.c_int),
                ("fx", C.POINTER(C.c_double)), ("fy", C.POINTER(C.c_double)), ("fz", C.POINTER(C.c_double))]


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
