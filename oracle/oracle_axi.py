"""ctypes binding of the axisymmetric CPU oracle (oracle/cfd_axi.c).

TEST INFRASTRUCTURE ONLY: the checker of the HIP axisymmetric solver and the
CPU baseline of its benchmark. The product never loads it.
"""
import ctypes as C

import numpy as np

import oracle3d

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
ZONE_F32 = ("phi", "ae", "an", "wgt", "lmix", "tgt", "rate", "yrate", "mfd")
_done = False


def lib():
    global _done
    L = oracle3d.lib()
    if not _done:
        L.ax_create.restype = C.c_void_p
        L.ax_create.argtypes = [C.c_int, C.c_int, f64p, f64p, u8p, f32p, f32p, f32p, u8p] + [f32p] * 6 + [f64p, f64p, f64p]
        L.ax_destroy.argtypes = [C.c_void_p]
        L.ax_set_opt.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double]
        L.ax_set_comps.argtypes = [C.c_void_p, f64p]
        L.ax_step.argtypes = [C.c_void_p, C.c_int]
        L.ax_read.argtypes = [C.c_void_p, f64p, f64p, f64p, f64p]
        L.ax_write.argtypes = [C.c_void_p, f64p]
        _done = True
    return L


class OracleAxi:
    """CPUSolver (src/cfd_cpu.js:50-210) restated in C."""

    def __init__(self, W, H, xf, rf, zones, comps, fs, Q0, opt):
        L = lib()
        self.W, self.H, self.N = W, H, W * H
        f = lambda a, dt: np.ascontiguousarray(a, dtype=dt)
        self._keep = [f(xf, np.float64), f(rf, np.float64), f(zones["solid"], np.uint8), f(zones["phi"], np.float32),
                      f(zones["ae"], np.float32), f(zones["an"], np.float32), f(zones["comp"], np.uint8)] + \
                     [f(zones[k], np.float32) for k in ("wgt", "lmix", "tgt", "rate", "yrate", "mfd")] + \
                     [f(comps, np.float64), np.array([fs["r"], fs["u"], fs["p"]], np.float64), f(np.concatenate(Q0), np.float64)]
        self.h = L.ax_create(W, H, *self._keep)
        self.set_opt(opt)

    def set_opt(self, opt):
        lib().ax_set_opt(self.h, opt.get("cfl", 0.6), opt.get("order", 2), int(opt.get("visc", True)), opt.get("hf", 0.0),
                         int(bool(opt.get("global", False))), float(opt.get("dtRatio", 0) or 0))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ax_destroy(self.h)
            self.h = None

    def set_comps(self, c):
        lib().ax_set_comps(self.h, np.ascontiguousarray(c, np.float64))

    def step(self, n):
        lib().ax_step(self.h, n)

    def read(self):
        Q = np.empty(5 * self.N); src = np.empty(self.N); mfx = np.empty(self.N); dt = np.empty(self.N)
        lib().ax_read(self.h, Q, src, mfx, dt)
        return Q.reshape(5, self.N), src, mfx, dt

    def write(self, Q):
        lib().ax_write(self.h, np.ascontiguousarray(np.asarray(Q, np.float64).reshape(-1)))
