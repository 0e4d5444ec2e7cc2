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
        ci, cd = C.c_int, C.c_double
        i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
        L.oax_thrust.argtypes = [ci, ci, f64p, f64p, f64p, u8p, f32p, f32p, u8p, f32p, f64p, f64p, f64p, f64p, cd, f64p]
        L.oax_section.argtypes = [ci, ci, f64p, f64p, f32p, f64p, f64p, ci, u8p, f64p]
        L.oax_derived.argtypes = [ci, ci, f64p, f64p, u8p, f32p, i8p, u8p, f64p, cd, cd, ci, f32p, u8p, u8p]
        L.oax_interp.argtypes = [ci, ci, f64p, f64p, u8p, f64p, ci, ci, f64p, f64p, cd, cd, f64p]
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


# ---------------------------------------------------------------------------
# post-processing (oracle/axi_post.c): cfdForces, cfdSection, cfdDerived,
# interpState on a state given as 5 planes (any float type, taken as doubles)
# ---------------------------------------------------------------------------
DERIVED_KEYS = ["T", "Tt", "P", "Pt", "V", "M", "rho", "Vx", "Vth", "Vr", "Y", "vort", "ds"]
SECTION_SUMS = ["mdot", "w", "A", "pA", "Tt", "Pt", "T", "P", "V", "M", "Y"]


def grid_arrays(xf, rf):
    """Cell centres and widths in doubles as buildCFDGrid forms them."""
    xf, rf = np.asarray(xf, np.float64), np.asarray(rf, np.float64)
    return {"W": len(xf) - 1, "H": len(rf) - 1, "xc": 0.5 * (xf[:-1] + xf[1:]), "dx": xf[1:] - xf[:-1],
            "rc": 0.5 * (rf[:-1] + rf[1:]), "dr": rf[1:] - rf[:-1]}


def _d(a):
    return np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1))


def thrust(g, zones, Q, src, mfx, comps, P0):
    """cfdForces with the solver's sources (src/cfd_post.js:44-75): F, Fb, Fw, mfuel, pw."""
    out = np.zeros(11)
    f = lambda a, dt: np.ascontiguousarray(a, dt)
    lib().oax_thrust(g["W"], g["H"], g["rc"], g["dr"], g["dx"], f(zones["solid"], np.uint8), f(zones["phi"], np.float32),
                     f(zones["ae"], np.float32), f(zones["comp"], np.uint8), f(zones["rate"], np.float32), _d(Q), _d(src), _d(mfx),
                     _d(comps), P0, out)
    return {"F": out[0] + out[1], "Fb": out[0], "Fw": out[1], "mfuel": out[2], "pw": out[3:].tolist()}


def section_sums(g, zones, Q, mfx, col, rows):
    """The 11 sums of cfdSection at column col over the row mask rows (SECTION_SUMS order)."""
    out = np.zeros(11)
    lib().oax_section(g["W"], g["H"], g["rc"], g["dr"], np.ascontiguousarray(zones["ae"], np.float32), _d(Q), _d(mfx), int(col),
                      np.ascontiguousarray(rows, np.uint8), out)
    return out


def derived(g, zones, Q, T0, P0, disp):
    """cfdDerived (src/cfd_post.js:10-40): (fields 13 x N float32 in DERIVED_KEYS order, reg, mat)."""
    N = g["W"] * g["H"]
    fields = np.zeros(13 * N, np.float32); reg = np.zeros(N, np.uint8); mat = np.zeros(N, np.uint8)
    lib().oax_derived(g["W"], g["H"], g["xc"], g["rc"], np.ascontiguousarray(zones["solid"], np.uint8),
                      np.ascontiguousarray(zones["phi"], np.float32), np.ascontiguousarray(zones["stream"], np.int8),
                      np.ascontiguousarray(zones["mat"], np.uint8), _d(Q), T0, P0, int(disp), fields, reg, mat)
    return fields.reshape(13, N), reg, mat


def interp(gA, solidA, QA, gB, T0, P0):
    """interpState (src/app_cfd.js:321-358): the state of grid A on grid B, 5 x NB doubles."""
    NB = gB["W"] * gB["H"]
    out = np.zeros(5 * NB)
    lib().oax_interp(gA["W"], gA["H"], gA["xc"], gA["rc"], np.ascontiguousarray(solidA, np.uint8), _d(QA), gB["W"], gB["H"],
                     gB["xc"], gB["rc"], T0, P0, out)
    return out.reshape(5, NB)
