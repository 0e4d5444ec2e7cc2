"""Per-cell comparison convention of the reference's GPU/CPU check
(tools/gpucmp.mjs:19-21): max |a - b| over fluid cells divided by the
per-variable scale max |b| over fluid cells."""
import numpy as np

VARS = ("r", "ru", "rv", "rw", "E")


def rel_diff(Q, Qref, t):
    fluid = t == 0
    out = {}
    for c, v in enumerate(VARS):
        a = Q[c][fluid].astype(np.float64)
        b = Qref[c][fluid].astype(np.float64)
        scale = max(np.abs(b).max(), 1e-30)
        out[v] = float(np.abs(a - b).max() / scale)
    return out


def rel(a, b, floor=1e-12):
    return abs(a - b) / max(abs(b), floor)
