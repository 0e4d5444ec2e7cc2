"""Axisymmetric oracle (oracle/cfd_axi.c) and JS host post-processing against
the reference's own outputs (tests/golden/axi_*).

The oracle restates CPUSolver (src/cfd_cpu.js) with the same double
arithmetic; it is not bit-exact because V8's Math.pow / Math.expm1 and glibc's
pow / expm1 differ in the last bit on some inputs (measured: pow on 9 % of
random pairs), which enters through the Riemann farfield; the bar is 1e-12
relative per variable after 500 iterations. The JS post functions (thrust,
section average, exit-temperature control) run on the reference's states and
must reproduce its numbers exactly, and so must the display fields
(cfdDerived).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_axi as OA
from axi_case import load, rel_diff
from golden_io import Golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


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


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
@pytest.mark.parametrize("case,it", [("axi_tj_xcoarse", 500), ("axi_hbtf_xcoarse", 100), ("axi_lbtf_xcoarse", 100)])
def test_js_post_bit_exact_on_reference_states(case, it):
    g, *_ = load(case)
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "axi_post_check.js"), g.dir, f"it{it}"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    h = [x for x in g.s["hist"] if x["it"] == it][0]
    for k in ("F", "Fb", "Fw", "mfuel", "pw"):
        assert out["thrust"][k] == h[k], k
    assert out["section"] == g.s["section"]
    assert out["control"]["comps_after"] == g.s["control"]["comps_after"]
    assert out["control"]["Tt4"] == g.s["control"]["Tt4"] and out["control"]["ratios"] == g.s["control"]["ratios"]
    # display fields (cfdDerived, src/cfd_post.js:10-40), both solid rules: every array byte-identical
    assert out["derived"] and all(out["derived"].values()), [k for k, v in out["derived"].items() if not v]
