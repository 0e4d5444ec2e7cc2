"""Shared comparisons of the engine field model (device or CPU build of
3d_aero_lab_amd/csrc/field.hpp) with the reference's own outputs
(tests/golden/axi_field, tests/golden/applycond, the setups' initial states)."""
import hashlib

import numpy as np

from golden_io import Golden

MODELS = ["tj", "hbtf", "lbtf", "tj_ab", "lbtf_ab"]
Q0_CASES = ["axi_tj_xcoarse", "axi_tj_coarse", "axi_tj_medium", "axi_hbtf_xcoarse", "axi_hbtf_medium", "axi_lbtf_xcoarse"]


def rel_rows(out, ref):
    """Per-value relative difference (0 where both are NaN, NaN against a
    number counts as a mismatch of 1)."""
    both = np.isnan(out) & np.isnan(ref)
    one = np.isnan(out) ^ np.isnan(ref)
    d = np.where(both, 0.0, np.abs(out - ref) / np.maximum(np.abs(ref), 1e-300))
    return np.where(one, 1.0, d)


def compare_samples(sample_fn):
    """sample_fn(desc, x, r, th) -> (n, 14). Per model: region mismatches and
    the largest relative difference of the 13 values."""
    g = Golden("axi_field")
    res = {}
    for name in MODELS:
        x, r, th = g.arr(f"{name}_x"), g.arr(f"{name}_r"), g.arr(f"{name}_th")
        ref = g.arr(f"{name}_out").reshape(-1, 14)
        out = sample_fn(g.s["models"][name], x, r, th)
        res[name] = {"n": int(x.size), "region_mismatch": int((out[:, 0] != ref[:, 0]).sum()),
                     "max_rel": float(rel_rows(out[:, 1:], ref[:, 1:]).max()),
                     "regions": sorted(set(int(v) for v in ref[:, 0]))}
    return res


def compare_initial_states(init_fn):
    """init_fn(desc, xf, rf, solid, stream, uniform) -> (5, N). Per setup with
    the reference's full initial state: the largest relative difference per
    cell, and whether the uniform start is bit-identical to the reference's."""
    fg = Golden("axi_field")
    res = {}
    for case in Q0_CASES:
        G = Golden(case)
        key = case[4:]
        desc = fg.s["models"][G.s["type"]]
        xf, rf = G.arr("xf"), G.arr("rf")
        solid, stream = G.arr("z_solid"), G.arr("z_stream")
        ref = np.stack([G.arr(f"Q0_{q}") for q in range(5)])
        Q = init_fn(desc, xf, rf, solid, stream, False)
        U = init_fn(desc, xf, rf, solid, stream, True)
        want_u = fg.s["initial_state"][key]["uniform_sha256"]
        res[case] = {"max_rel": float(rel_rows(Q, ref).max()),
                     "cells_differing": int((Q != ref).any(axis=0).sum()), "cells": int(Q.shape[1]),
                     "uniform_exact": all(hashlib.sha256(np.ascontiguousarray(U[q]).tobytes()).hexdigest() == want_u[q] for q in range(5))}
    return res
