"""ctypes binding of the CPU oracle (oracle/cfd3d.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker and by bench.py's
cpu_baseline leg as the CPU baseline. The product never loads it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle3d.so")
_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.o3_create.restype = C.c_void_p
        L.o3_create.argtypes = [C.c_int] * 3 + [f32p] * 3 + [u8p] + [f32p] * 5
        L.o3_destroy.argtypes = [C.c_void_p]
        L.o3_set_free.argtypes = [C.c_void_p, f64p]
        L.o3_set_fix.argtypes = [C.c_void_p, f64p]
        L.o3_set_opt.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.o3_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.o3_iter.argtypes = [C.c_void_p]
        L.o3_reset.argtypes = [C.c_void_p]
        L.o3_step.argtypes = [C.c_void_p, C.c_int]
        L.o3_read.argtypes = [C.c_void_p] + [f32p] * 5
        L.o3_write.argtypes = [C.c_void_p] + [f32p] * 5
        L.o3_read_dt.argtypes = [C.c_void_p, f64p]
        L.o3_forces.argtypes = [f32p] * 5 + [i32p] + [f64p] * 4 + [C.c_int] + [C.c_double] * 6 + [f64p]
        L.o3_trefftz.restype = C.c_double
        L.o3_trefftz.argtypes = [C.c_int] * 3 + [f32p] * 3 + [u8p] + [f32p] * 4 + [C.c_double] * 4
        L.o3_residual.restype = C.c_double
        L.o3_residual.argtypes = [C.c_int, u8p, f32p, f32p, C.c_int]
        L.o3_mach_max.restype = C.c_double
        L.o3_mach_max.argtypes = [C.c_int, u8p] + [f32p] * 5
        _lib = L
    return _lib


class Oracle3D:
    """CfdCPU restated in C: same protocol (set_free, set_fix, reset, step, read)."""

    def __init__(self, grid, geo, threads=1):
        L = lib()
        self.NX, self.NY, self.NZ = grid["NX"], grid["NY"], grid["NZ"]
        self.N = self.NX * self.NY * self.NZ
        a = lambda x, dt: np.ascontiguousarray(x, dtype=dt)
        self._keep = [a(grid["dx"], np.float32), a(grid["dy"], np.float32), a(grid["dz"], np.float32),
                      a(geo["t"], np.uint8), a(geo["phi"], np.float32)] + [a(geo["ap"][d], np.float32) for d in range(4)]
        self.h = L.o3_create(self.NX, self.NY, self.NZ, *self._keep)
        if not self.h:
            raise ValueError(f"grid {self.NX}x{self.NY}x{self.NZ} out of range (NX, NZ >= 1, NY >= 2)")
        L.o3_set_threads(self.h, threads)

    def __del__(self):
        if getattr(self, "h", None):
            lib().o3_destroy(self.h)
            self.h = None

    def set_free(self, F):
        lib().o3_set_free(self.h, np.asarray(F, dtype=np.float64))

    def set_fix(self, fix):
        lib().o3_set_fix(self.h, np.ascontiguousarray(np.asarray(fix, dtype=np.float64).reshape(15)))

    def set_opt(self, cfl, order=2):
        lib().o3_set_opt(self.h, cfl, order)

    def set_threads(self, n):
        lib().o3_set_threads(self.h, n)

    def reset(self):
        lib().o3_reset(self.h)

    def step(self, n):
        lib().o3_step(self.h, n)

    @property
    def it(self):
        return lib().o3_iter(self.h)

    def read(self):
        out = [np.empty(self.N, np.float32) for _ in range(5)]
        lib().o3_read(self.h, *out)
        return np.stack(out)

    def write(self, Q):
        lib().o3_write(self.h, *[np.ascontiguousarray(Q[c], dtype=np.float32) for c in range(5)])

    def read_dt(self):
        out = np.empty(self.N, np.float64)
        lib().o3_read_dt(self.h, out)
        return out


def forces(Q, fl, alpha, M, S, xcg, zcg, cref):
    out = np.zeros(3, np.float64)
    lib().o3_forces(*[np.ascontiguousarray(Q[c], dtype=np.float32) for c in range(5)],
                    np.ascontiguousarray(fl["q"], dtype=np.int32), *[np.ascontiguousarray(fl[k], dtype=np.float64) for k in ("cx", "cz", "x", "z")],
                    len(fl["q"]), alpha, M, S, xcg, zcg, cref, out)
    return {"CL": out[0], "CD": out[1], "Cm": out[2]}


def trefftz(grid, t, Q, xT, alpha, M, S):
    a = lambda x: np.ascontiguousarray(x, dtype=np.float32)
    return lib().o3_trefftz(grid["NX"], grid["NY"], grid["NZ"], a(grid["xc"]), a(grid["dy"]), a(grid["dz"]),
                            np.ascontiguousarray(t, dtype=np.uint8), a(Q[0]), a(Q[1]), a(Q[2]), a(Q[3]), xT, alpha, M, S)


def residual(t, r, prev, nit):
    return lib().o3_residual(len(r), np.ascontiguousarray(t, dtype=np.uint8), np.ascontiguousarray(r, dtype=np.float32),
                             np.ascontiguousarray(prev, dtype=np.float32), nit)


def mach_max(t, Q):
    a = lambda x: np.ascontiguousarray(x, dtype=np.float32)
    return lib().o3_mach_max(len(Q[0]), np.ascontiguousarray(t, dtype=np.uint8), *[a(Q[c]) for c in range(5)])
