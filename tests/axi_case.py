"""Loads an axisymmetric golden case (setup inputs of the reference CPUSolver)."""
import numpy as np

from golden_io import Golden

ZONE_KEYS = ("solid", "phi", "ae", "an", "comp", "wgt", "lmix", "tgt", "rate", "yrate", "mfd", "stream")


def load(case):
    g = Golden(case)
    z = {k: g.arr("z_" + k) for k in ZONE_KEYS}
    Q0 = np.stack([g.arr(f"Q0_{q}") for q in range(5)])
    return g, z, Q0, g.arr("comps"), g.arr("xf"), g.arr("rf")


def rel_diff(Q, Qref, solid):
    fl = solid == 0
    return {f"Q{q}": float(np.abs(Q[q][fl].astype(np.float64) - Qref[q][fl]).max() / np.abs(Qref[q][fl]).max()) for q in range(5)}
