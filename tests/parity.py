"""Per-cell comparison convention of the reference's GPU/CPU check
(tools/gpucmp.mjs:19-21): max |a - b| over fluid cells divided by the
per-variable scale max |b| over fluid cells."""
import numpy as np

VARS = ("r", "ru", "rv", "rw", "E")


def rel_diff(Q, Qref, t, floor=1e-30):
    """`floor` bounds the scale from below: on a grid with a handful of fluid
    cells a variable can stay at round-off level (rho v ~ 1e-9 with no sideslip),
    where the reference's convention would divide by that round-off."""
    fluid = t == 0
    out = {}
    for c, v in enumerate(VARS):
        a = Q[c][fluid].astype(np.float64)
        b = Qref[c][fluid].astype(np.float64)
        if a.size == 0:
            out[v] = 0.0
            continue
        scale = max(np.abs(b).max(), floor)
        out[v] = float(np.abs(a - b).max() / scale)
    return out


def rel(a, b, floor=1e-12):
    return abs(a - b) / max(abs(b), floor)
