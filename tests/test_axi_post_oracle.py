"""The axisymmetric post-processing oracle (oracle/axi_post.c) against the
reference's own outputs on its own states (tests/golden/axi_*, written by
tests/golden/make_golden_axi.js from the reference's CPUSolver):
  thrust   cfdForces with the solver's sources (src/cfd_post.js:44-75)
  section  cfdSection at the combustor exit (src/cfd_post.js:83-102)
  derived  cfdDerived's display fields, both solid rules (src/cfd_post.js:10-40)
  interp   interpState onto the medium grid (src/app_cfd.js:321-358)
and the JS exit-temperature control (js/cfd_post.js controlFromSections) on
the reference's section.

Bars: every value that is sums, products and quotients of the state equals
the reference's bit for bit; values through Math.pow / Math.log (section Pt,
display Pt and ds) within a few ulps, because V8's pow / log and glibc's
differ in the last bit on some inputs (display fields, stored in FP32: within
one FP32 ulp of the reference's).
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_axi as OA
from axi_case import load
from golden_io import Golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = [("axi_tj_xcoarse", (10, 100, 500)), ("axi_hbtf_xcoarse", (10, 100)), ("axi_lbtf_xcoarse", (10, 100))]


def col_at(xc, x):
    """Column whose centre is nearest to x; a tie goes to the right one."""
    d = np.abs(np.asarray(xc) - x)
    return int(np.flatnonzero(d == d.min())[-1])


def station_rows(z, W, H, col, si):
    k = np.arange(H) * W + col
    return ((z["solid"][k] == 0) & (z["stream"][k] == si)).astype(np.uint8)


def averages(sums, x):
    """cfdSection's averages from the oracle's sums (division by w, or 1 with no positive flow)."""
    s = dict(zip(OA.SECTION_SUMS, sums))
    d = s["w"] or 1.0
    out = {"x": x, "mdot": s["mdot"], "A": s["A"]}
    for k in ("Tt", "Pt", "T", "P", "V", "M", "Y"):
        out[k] = s[k] / d
    if not s["w"]:
        out["P"] = s["pA"] / (s["A"] or 1.0)
    return out


def control_section(g, z):
    s = g.s
    grid = OA.grid_arrays(g.arr("xf"), g.arr("rf"))
    x = s["ctl"]["comb"]["x1"] - 0.004 * s["ctl"]["s"]
    col = col_at(grid["xc"], x)
    rows = station_rows(z, grid["W"], grid["H"], col, s["ctl"]["order"].index(s["ctl"]["comb"]["stream"]))
    return grid, col, rows


@pytest.mark.parametrize("case,tags", CASES)
def test_oracle_thrust_equals_reference(case, tags):
    g, z, Q0, comps, xf, rf = load(case)
    grid = OA.grid_arrays(xf, rf)
    hist = {h["it"]: h for h in g.s["hist"]}
    for it in tags:
        Q = np.stack([g.arr(f"it{it}_Q{q}") for q in range(5)])
        t = OA.thrust(grid, z, Q, g.arr(f"it{it}_src"), g.arr(f"it{it}_mfx"), comps, g.s["ctl"]["P0"])
        for k in ("F", "Fb", "Fw", "mfuel", "pw"):
            assert t[k] == hist[it][k], (case, it, k, t[k], hist[it][k])


@pytest.mark.parametrize("case,tags", CASES)
def test_oracle_section_equals_reference(case, tags):
    g, z, *_ = load(case)
    it = tags[-1]
    Q = np.stack([g.arr(f"it{it}_Q{q}") for q in range(5)])
    grid, col, rows = control_section(g, z)
    got = averages(OA.section_sums(grid, z, Q, g.arr(f"it{it}_mfx"), col, rows), float(grid["xc"][col]))
    ref = g.s["section"]
    assert set(got) == set(ref)
    for k in ref:
        if k == "Pt":
            assert abs(got[k] - ref[k]) <= 4e-16 * abs(ref[k]), (k, got[k], ref[k])
        else:
            assert got[k] == ref[k], (k, got[k], ref[k])


@pytest.mark.parametrize("case,tags", CASES)
def test_oracle_derived_equals_reference(case, tags):
    g, z, *_ = load(case)
    z["mat"] = g.arr("z_mat")
    it = tags[-1]
    Q = np.stack([g.arr(f"it{it}_Q{q}") for q in range(5)])
    grid = OA.grid_arrays(g.arr("xf"), g.arr("rf"))
    for disp in (0, 1):
        fields, reg, mat = OA.derived(grid, z, Q, g.s["fl"]["T0"], g.s["fl"]["P0"], disp)
        assert np.array_equal(reg, g.arr(f"der{disp}_reg")) and np.array_equal(mat, g.arr(f"der{disp}_mat"))
        for c, k in enumerate(OA.DERIVED_KEYS):
            ref = g.arr(f"der{disp}_{k}")
            if k in ("Pt", "ds"):
                fin = np.isfinite(ref)
                assert np.array_equal(fin, np.isfinite(fields[c]))
                ulp = np.spacing(np.abs(ref[fin])).astype(np.float64)
                assert (np.abs(fields[c][fin].astype(np.float64) - ref[fin]) <= ulp).all(), (case, disp, k)
            else:
                assert np.array_equal(fields[c].view(np.uint32), ref.view(np.uint32)), (case, disp, k)


def test_oracle_interp_equals_reference():
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "axi_interp", "manifest.json")))["interp"]
    import hashlib
    assert set(ref) == {"tj", "hbtf"}
    for typ, r in ref.items():
        A, B = Golden(f"axi_{typ}_{r['from']}"), Golden(f"axi_{typ}_{r['to']}")
        gA, gB = OA.grid_arrays(A.arr("xf"), A.arr("rf")), OA.grid_arrays(B.arr("xf"), B.arr("rf"))
        QA = np.stack([A.arr(f"it{r['iters']}_Q{q}") for q in range(5)])
        QB = OA.interp(gA, A.arr("z_solid"), QA, gB, B.s["fl"]["T0"], B.s["fl"]["P0"])
        for q in range(5):
            assert hashlib.sha256(QB[q].tobytes()).hexdigest() == r["sha256"][q], (typ, q)
            assert hashlib.sha256(QB[q].astype(np.float32).tobytes()).hexdigest() == r["sha256_f32"][q], (typ, q)


def test_oracle_section_is_station_sum():
    """A station average (tools/stations.js:8-16) is the same sums over its
    stream's rows: closed forms on a synthetic state (uniform flow, one
    column), so a dropped weight or a wrong area fails."""
    W, H = 4, 3
    grid = OA.grid_arrays(np.array([0.0, 1, 2, 3, 4]), np.array([0.0, 0.5, 1.5, 2.0]))
    z = {"ae": np.ones(W * H, np.float32)}
    r, u, p = 1.2, 50.0, 1e5
    g = 1.4
    Q = np.zeros((5, W * H))
    Q[0], Q[1], Q[3] = r, r * u, p / (g - 1) + 0.5 * r * u * u
    mfx = np.full(W * H, r * u)
    rows = np.array([1, 0, 1], np.uint8)
    s = dict(zip(OA.SECTION_SUMS, OA.section_sums(grid, z, Q, mfx, 2, rows)))
    area = 2 * np.pi * (grid["rc"] * grid["dr"])[[0, 2]]
    np.testing.assert_allclose(s["mdot"], (r * u * area).sum(), rtol=1e-14)
    np.testing.assert_allclose(s["A"], area.sum(), rtol=1e-14)
    T = p / (r * 287.05)
    cp = g * 287.05 / (g - 1)
    Tt = T + u * u / (2 * cp)
    np.testing.assert_allclose(s["Tt"] / s["w"], Tt, rtol=1e-14)
    np.testing.assert_allclose(s["Pt"] / s["w"], p * (Tt / T) ** (g / (g - 1)), rtol=1e-13)
    np.testing.assert_allclose(s["M"] / s["w"], u / np.sqrt(g * p / r), rtol=1e-14)


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
@pytest.mark.parametrize("case,tags", CASES)
def test_js_control_on_reference_section(case, tags):
    """controlFromSections (the exit-temperature control and spool ratios of
    cfdControl, src/cfd_post.js:104-126) fed the reference's combustor
    section and blade-row powers: the reference's gains, Tt4 and ratios."""
    g = Golden(case)
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "axi_control_check.js"), g.dir], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout)
    ref = g.s["control"]
    assert out["comps_after"] == ref["comps_after"]
    assert out["Tt4"] == ref["Tt4"] and out["ratios"] == ref["ratios"]
