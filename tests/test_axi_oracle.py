"""Axisymmetric oracle (oracle/cfd_axi.c) against the reference's own outputs
(tests/golden/axi_*).

The oracle restates CPUSolver (src/cfd_cpu.js) with the same double
arithmetic; it is not bit-exact because V8's Math.pow / Math.expm1 and glibc's
pow / expm1 differ in the last bit on some inputs (measured: pow on 9 % of
random pairs), which enters through the Riemann farfield; the bar is 1e-12
relative per variable after 500 iterations. The post-processing oracle is
pinned in test_axi_post_oracle.py.
"""
import numpy as np
import pytest

import oracle_axi as OA
from axi_case import load, rel_diff
from golden_io import Golden


@pytest.mark.parametrize("case,iters", [("axi_tj_xcoarse", (10, 100, 500)), ("axi_hbtf_xcoarse", (10, 100)),
                                        ("axi_lbtf_xcoarse", (10, 100))])
def test_axi_oracle_matches_reference(case, iters):
    """TJ (one stream), HBTF (inlet / core / bypass streams, fan and booster
    rows, split cells) and LBTF (fan, mixed exhaust)."""
    g, z, Q0, comps, xf, rf = load(case)
    s = g.s
    o = OA.OracleAxi(s["W"], s["H"], xf, rf, z, comps, s["fs"], list(Q0), dict(s["opt"], hf=s["hf"]))
    done = 0
    for it in iters:
        o.step(it - done)
        done = it
        Q, src, mfx, dt = o.read()
        ref = np.stack([g.arr(f"it{it}_Q{q}") for q in range(5)])
        d = rel_diff(Q, ref, z["solid"])
        assert max(d.values()) < 1e-12, (it, d)
        fl = z["solid"] == 0
        assert np.abs(mfx[fl] - g.arr(f"it{it}_mfx")[fl]).max() <= 1e-12 * np.abs(g.arr(f"it{it}_mfx")).max()


OPTION_RUNS = [("global", {"global": True}, [(10, None)]), ("dtr3", {"dtRatio": 3}, [(10, None)]),
               ("o1e", {"order": 1, "visc": False}, [(10, None)]), ("dtr35", {"dtRatio": 3}, [(30, None), (30, {"dtRatio": 5})])]


@pytest.mark.parametrize("tag,opt,sched", OPTION_RUNS)
def test_axi_oracle_option_paths(tag, opt, sched):
    """CPUSolver's option paths (src/cfd_cpu.js:53, 193-203) on TJ xcoarse
    against the reference's own runs: the global and dtRatio dt caps, first
    order without the viscous r fluxes, and a dtRatio change mid-run."""
    g0, z, Q0, comps, xf, rf = load("axi_tj_xcoarse")
    g = Golden("axi_tj_options")
    s = g0.s
    base = dict(s["opt"], hf=s["hf"])
    o = OA.OracleAxi(s["W"], s["H"], xf, rf, z, comps, s["fs"], list(Q0), dict(base, **opt))
    done = 0
    for n, change in sched:
        if change:
            o.set_opt({**base, **opt, **change})
        o.step(n)
        done += n
    Q, _, _, dt = o.read()
    ref = np.stack([g.arr(f"{tag}_it{done}_Q{q}") for q in range(5)])
    d = rel_diff(Q, ref, z["solid"])
    assert max(d.values()) < 1e-12, d
    if tag == "global":   # one dt for every fluid cell
        fl = z["solid"] == 0
        assert np.abs(dt[fl] - g.arr("global_it10_dt")[fl]).max() <= 1e-12 * g.arr("global_it10_dt")[fl].max()
        assert np.unique(dt[fl]).size == 1


