"""Runs the CPU build of the tiled stage kernel against the oracle.

Usage (needs libasan preloaded, see tests/test_tile_emul.py):
    python tests/cpu_emul/run_emul.py <lib> <case> <iters> <ty> <zc>
<case> is a golden case or synth:NX:NY:NZ (tests/synth_cases.py, seed NX+NY+NZ).
Prints the per-variable relative difference as JSON.
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import oracle3d as O  # noqa: E402
from golden_io import Golden  # noqa: E402
from parity import rel_diff  # noqa: E402
from synth_cases import synth_case  # noqa: E402


def prim(a):
    r = max(a[0], 1e-6); u, v, w = a[1] / r, a[2] / r, a[3] / r
    return [r, u, v, w, max(0.3999999999999999 * (a[4] - 0.5 * r * (u * u + v * v + w * w)), 1e-7)]


def main():
    lib, case, iters, ty, zc = sys.argv[1], sys.argv[2], int(sys.argv[3]), int(sys.argv[4]), int(sys.argv[5])
    L = C.CDLL(lib)
    if case.startswith("synth:"):
        # synth:NX:NY:NZ[@mach]
        dims, _, mach = case.partition("@")
        n = [int(x) for x in dims.split(":")[1:]]
        grid, geo, F, fix = synth_case(*n, seed=sum(n), mach=float(mach) if mach else 0.6)
        cfl = 0.5
    else:
        g = Golden(case)
        grid, geo = g.grid(), g.geo()
        F = g.s["F"]
        fix = g.s.get("fix", [F, F, F])
        cfl = g.s["cfl"]
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    N = grid["NX"] * grid["NY"] * grid["NZ"]
    ap = f32(np.concatenate(geo["ap"]))
    Q = f32(np.stack([np.full(N, F[c], np.float32) for c in range(5)]).reshape(-1))
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    Fw = np.array(prim(F), np.float64)
    fixw = np.array([prim(f) for f in fix], np.float64).reshape(-1)
    fixq = np.array(fix, np.float64).reshape(-1)
    keep = [f32(grid["dx"]), f32(grid["dy"]), f32(grid["dz"]), np.ascontiguousarray(geo["t"], np.uint8), f32(geo["phi"])]
    L.emul_steps.argtypes = [C.c_int] * 3 + [C.c_void_p] * 9 + [C.c_double, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int]
    rc = L.emul_steps(grid["NX"], grid["NY"], grid["NZ"], *[P(k) for k in keep], P(ap), P(Fw), P(fixw), P(fixq),
                      cfl, 2, P(Q), iters, ty, zc)
    assert rc == 0
    o = O.Oracle3D(grid, geo)
    o.set_free(F); o.set_fix(fix); o.set_opt(cfl, 2); o.reset(); o.step(iters)
    # synthetic grids: scale floor of 1e-3 of the freestream's largest component
    d = rel_diff(Q.reshape(5, N), o.read(), geo["t"], floor=1e-3 * max(abs(f) for f in F) if case.startswith("synth:") else 1e-30)
    print(json.dumps(d))


if __name__ == "__main__":
    main()
