"""CPU build of the axisymmetric kernels against the axisymmetric oracle.
    python tests/cpu_emul/run_axi_emul.py <lib> <case> <iters> [fused]   -> JSON of relative differences
(and, with `fused`, whether the fused tile kernel gave the same bits as the face + cell kernels)"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import oracle_axi as OA  # noqa: E402
from golden_io import Golden  # noqa: E402

ZF = ("phi", "ae", "an", "wgt", "lmix", "tgt", "rate", "yrate", "mfd")


def main():
    lib, case, iters = sys.argv[1], sys.argv[2], int(sys.argv[3])
    fused = len(sys.argv) > 4 and sys.argv[4] == "fused"
    g = Golden(case)
    s = g.s
    W, H = s["W"], s["H"]
    N = W * H
    z = {k: g.arr("z_" + k) for k in ("solid", "comp") + ZF}
    Q0 = np.stack([g.arr(f"Q0_{q}") for q in range(5)])
    comps = g.arr("comps")
    fs = np.array([s["fs"]["r"], s["fs"]["u"], s["fs"]["p"]], np.float64)
    L = C.CDLL(lib)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    L.axi_emul_steps.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 7 + [C.c_double, C.c_int, C.c_int, C.c_double] + [C.c_void_p] * 3 + [C.c_int, C.c_int]
    xf, rf = g.arr("xf"), g.arr("rf")
    zf = np.ascontiguousarray(np.concatenate([z[k] for k in ZF]), np.float32)
    Q = np.ascontiguousarray(Q0.astype(np.float32).reshape(-1))
    src, mfx = np.zeros(N, np.float32), np.zeros(N, np.float32)
    solid, comp = np.ascontiguousarray(z["solid"], np.uint8), np.ascontiguousarray(z["comp"], np.uint8)
    Q2 = Q.copy()
    L.axi_emul_steps(W, H, P(xf), P(rf), P(solid), P(comp), P(zf), P(comps), P(fs), s["opt"]["cfl"], 2, 1, s["hf"], P(Q), P(src), P(mfx), iters, 0)
    same = None
    if fused:
        same = True
        for form in (1, 2):          # two tile shapes, one with fewer waves than tile rows
            Q3 = Q2.copy()
            src2, mfx2 = np.zeros(N, np.float32), np.zeros(N, np.float32)
            L.axi_emul_steps(W, H, P(xf), P(rf), P(solid), P(comp), P(zf), P(comps), P(fs), s["opt"]["cfl"], 2, 1, s["hf"], P(Q3), P(src2),
                             P(mfx2), iters, form)
            same = same and bool(np.array_equal(Q.view(np.uint32), Q3.view(np.uint32)) and np.array_equal(mfx.view(np.uint32), mfx2.view(np.uint32)))
    o = OA.OracleAxi(W, H, xf, rf, z, comps, s["fs"], list(Q0), dict(s["opt"], hf=s["hf"]))
    o.step(iters)
    Qo, so, mo, _ = o.read()
    fl = z["solid"] == 0
    d = {f"Q{q}": float(np.abs(Q.reshape(5, N)[q][fl] - Qo[q][fl]).max() / np.abs(Qo[q][fl]).max()) for q in range(5)}
    d["mfx"] = float(np.abs(mfx[fl] - mo[fl]).max() / np.abs(mo[fl]).max())
    if same is not None:
        d["fused_same"] = 0.0 if same else 1.0
    print(json.dumps(d))


if __name__ == "__main__":
    main()
