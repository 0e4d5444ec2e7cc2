"""aero3d_readback_static (the readback with no device allocation, over
caller-owned device arrays and a static scratch buffer) against the handle's
aero3d_readback on the same states: residual, coefficients, Trefftz CDi and
Mach maximum bit for bit, on two calls in a row, the second on a smaller
field. The caller's device arrays come from the HIP runtime the library
itself links (libamdhip64 through ctypes), standing in for a caller's own
device memory."""
import ctypes as C
import sys

import numpy as np
import pytest

from golden_io import Golden

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-no-alloc]", *a, flush=True)
    sys.stdout.flush()


class DeviceArrays:
    """Caller-side device copies of host arrays (hipMalloc / hipMemcpy of the
    library's own runtime); freed together."""

    def __init__(self):
        import aero3d
        aero3d.lib()                                        # loads libamdhip64 with the library
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        return p.value

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def test_static_readback_equals_handle_readback():
    import aero3d
    D = DeviceArrays()
    for case, cfl in (("transport60k", 0.5), ("sphere", 0.5)):    # the second field is smaller
        g = Golden(case)
        F = g.s["F"]
        fix = g.s.get("fix", [F, F, F])
        grid, geo, fl = g.grid(), g.geo(), g.force_list()
        r = g.s.get("ref") or {"S": g.s["S"], "xcg": 0.0, "c": 1.0}
        alpha = g.s.get("alpha", 0.0)
        ref = (alpha, g.s["M"], r["S"], r["xcg"], 0.0, r["c"])
        xT = g.s.get("xT", float(grid["xc"][len(grid["xc"]) * 3 // 4]))
        s = aero3d.CfdHIP(grid, geo)
        s.opt["cfl"] = cfl
        s.set_free(F); s.set_fix(fix); s.reset(); s.set_force_list(fl)
        s.step(5)
        s.readback(*ref, xT)                         # snapshot at 5
        Qa = s.read()
        s.step(7)
        want = s.readback(*ref, xT)
        Qb = s.read()
        s.dispose()
        N = len(Qb[0])
        say(case, N, "cells; handle readback", want)
        Qd, snap, t = D.put(Qb), D.put(Qa[0]), D.put(geo["t"])
        dy, dz = D.put(grid["dy"]), D.put(grid["dz"])
        fq, cx, cz, x, z = (D.put(fl[k]) for k in ("q", "cx", "cz", "x", "z"))
        xc = np.ascontiguousarray(grid["xc"], np.float32)
        f = aero3d.Field(grid["NX"], grid["NY"], grid["NZ"], N, Qd, snap, 7, t, dy, dz, xc.ctypes.data, fq, cx, cz, x, z, len(fl["q"]))
        got = aero3d.readback_static(f, *ref, xT)
        say(case, "static readback", got)
        for k in ("res", "CL", "CD", "Cm", "CDi", "Mmax"):
            assert got[k] == want[k], (case, k, got[k], want[k])
        snap_h = D.get(snap, N, np.float32)
        assert np.array_equal(snap_h[::3], Qb[0][::3])          # the snapshot now holds the current density
        # a first call (nit 0) only takes the snapshot
        f.nit = 0
        assert np.isnan(aero3d.readback_static(f, *ref)["res"])
    # a field larger than the scratch supports is rejected before any launch
    big = aero3d.Field(1 << 11, 1 << 10, 1 << 10, 1 << 31, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0)
    with pytest.raises(aero3d.Aero3DError, match="bad field"):
        aero3d.readback_static(big, *ref)
    D.free()
