"""Axisymmetric post-processing on the device (axi_monitor, axi_sections,
axi_derived, axi_transfer) against the post-processing oracle
(oracle/axi_post.c, pinned on the reference's own outputs by
test_axi_post_oracle.py) on the same FP32 state read back from the device.

Bars: section sums (control sections, stations) and display fields bit for
bit, except the terms through pow / log (section Pt within 2e-15 relative;
display Pt and ds within one FP32 ulp); the thrust and the density residual
of the monitor within 1e-12 (grid-wide sums in another order); the transfer
onto another grid bit for bit with the oracle's result rounded to FP32.
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle_axi as OA
from axi_case import load
from golden_io import Golden

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STATION_TYPES = {"axi_tj_xcoarse": "tj", "axi_hbtf_xcoarse": "hbtf", "axi_lbtf_xcoarse": "lbtf", "axi_tj_medium": "tj"}
STATIONS = {   # tools/stations.js:8-16
    "hbtf": [("2", 0.58, "inlet"), ("13", 0.95, "bypass"), ("25", 1.47, "core"), ("3", 2.05, "core"), ("4", 2.43, "core"),
             ("45", 2.62, "core"), ("5", 3.08, "core"), ("19", 2.19, "bypass")],
    "tj": [("2", 0.36, "core"), ("3", 1.55, "core"), ("4", 2.21, "core"), ("5", 2.62, "core"), ("7", 4.29, "core"), ("9", 4.89, "core")],
    "lbtf": [("2", 0.3, "inlet"), ("13", 0.9, "bypass"), ("3", 1.55, "core"), ("4", 2.0, "core"), ("5", 2.55, "core"), ("6A", 3.0, "ab"),
             ("9", 4.94, "ab")],
}


def say(*a):
    print("[gpu-axi-post]", *a, flush=True)
    sys.stdout.flush()


def col_at(xc, x):
    d = np.abs(np.asarray(xc) - x)
    return int(np.flatnonzero(d == d.min())[-1])


def rows_of(z, W, H, col, si):
    k = np.arange(H) * W + col
    return ((z["solid"][k] == 0) & (z["stream"][k] == si)).astype(np.uint8)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("case,iters", [("axi_tj_xcoarse", 300), ("axi_hbtf_xcoarse", 200), ("axi_lbtf_xcoarse", 200),
                                        ("axi_tj_medium", 100)])
def test_axi_post_on_device_equals_oracle(case, iters):
    import aero3d
    g, z, Q0, comps, xf, rf = load(case)
    z["mat"] = g.arr("z_mat")
    s = g.s
    grid = OA.grid_arrays(xf, rf)
    W, H = grid["W"], grid["H"]
    ctl = s["ctl"]
    say(case, iters, "iterations, then 50 more")
    h = aero3d.AxiHIP(xf, rf, z, comps, s["fs"], Q0, dict(s["opt"], hf=s["hf"]))
    cx = ctl["comb"]["x1"] - 0.004 * ctl["s"]
    ccol = col_at(grid["xc"], cx)
    crow = rows_of(z, W, H, ccol, ctl["order"].index(ctl["comb"]["stream"]))
    h.monitor_sections([ccol], crow)
    h.step(iters)
    m1 = h.monitor(comps, ctl["P0"])
    rho1 = h.read()[0][0].copy()
    h.step(50)
    m2 = h.monitor(comps, ctl["P0"])
    Q, src, mfx, _ = h.read()
    assert not m1["res_valid"] and m2["res_valid"] and m2["bad"] == 0

    # thrust and blade-row powers (grid sums in another order)
    t = OA.thrust(grid, z, Q, src, mfx, comps, ctl["P0"])
    for k in ("F", "Fb", "Fw", "mfuel"):
        assert rel(m2[k], t[k]) < 1e-12, (k, m2[k], t[k])
    for q in range(8):
        if t["pw"][q] or m2["pw"][q]:
            assert rel(m2["pw"][q], t["pw"][q]) < 1e-12, (q, m2["pw"][q], t["pw"][q])
    # density residual over every second non-solid cell (src/app_cfd.js:368-371)
    k = np.arange(0, Q.shape[1], 2)
    k = k[z["solid"][k] == 0]
    d = (Q[0][k].astype(np.float64) - rho1[k]) / rho1[k]
    d = d[d == d]
    assert m2["res_n"] == d.size and rel(m2["res_sum2"], float((d * d).sum())) < 1e-12
    # the control section of the monitor, bit for bit
    o = OA.section_sums(grid, z, Q, mfx, ccol, crow)
    assert (m2["sec"][0]["mdot"], m2["sec"][0]["w"], m2["sec"][0]["wTt"]) == (o[0], o[1], o[4])

    # stations and the control section through axi_sections
    order = ctl["order"]
    secs = [(col_at(grid["xc"], x), rows_of(z, W, H, col_at(grid["xc"], x), order.index(sn))) for _, x, sn in STATIONS[STATION_TYPES[case]]]
    secs.append((ccol, crow))
    dev = h.sections([c for c, _ in secs], np.stack([r for _, r in secs]))
    for q, (c, r) in enumerate(secs):
        o = OA.section_sums(grid, z, Q, mfx, c, r)
        for i, name in enumerate(OA.SECTION_SUMS):
            if name == "Pt":
                assert rel(dev[q, i], o[i]) <= 2e-15, (q, name, dev[q, i], o[i])
            else:
                assert dev[q, i] == o[i], (q, name, dev[q, i], o[i])

    # display fields, both solid rules
    h.set_display_zones(z["stream"], z["mat"])
    for disp in (0, 1):
        f, reg, mat = h.derived(s["fl"]["T0"], s["fl"]["P0"], disp)
        fo, rego, mato = OA.derived(grid, z, Q, s["fl"]["T0"], s["fl"]["P0"], disp)
        assert np.array_equal(reg, rego) and np.array_equal(mat, mato)
        for c, key in enumerate(OA.DERIVED_KEYS):
            a, b = f[c], fo[c]
            if key in ("Pt", "ds"):
                fin = np.isfinite(b)
                assert np.array_equal(fin, np.isfinite(a)), key
                assert (np.abs(a[fin].astype(np.float64) - b[fin]) <= np.spacing(np.abs(b[fin])).astype(np.float64)).all(), (disp, key)
            else:
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (disp, key)
    say(case, "F", m2["F"], "station Tt", [round(float(dev[q, 4] / dev[q, 1]), 3) if dev[q, 1] else None for q in range(len(secs) - 1)])
    h.dispose()


@pytest.mark.parametrize("typ,iters", [("tj", 500), ("hbtf", 100)])
def test_axi_transfer_equals_oracle(typ, iters):
    """Grid sequencing on the device: the xcoarse state (the reference's own,
    at the fixture's iteration) carried onto the medium grid."""
    import aero3d
    gA, zA, Q0A, cA, xfA, rfA = load(f"axi_{typ}_xcoarse")
    gB, zB, Q0B, cB, xfB, rfB = load(f"axi_{typ}_medium")
    QA = np.stack([gA.arr(f"it{iters}_Q{q}") for q in range(5)])
    A = aero3d.AxiHIP(xfA, rfA, zA, cA, gA.s["fs"], QA, dict(gA.s["opt"], hf=gA.s["hf"]))
    B = aero3d.AxiHIP(xfB, rfB, zB, cB, gB.s["fs"], Q0B, dict(gB.s["opt"], hf=gB.s["hf"]))
    fl = gB.s["fl"]
    B.transfer_from(A, fl["T0"], fl["P0"])
    QB = B.read()[0]
    o = OA.interp(OA.grid_arrays(xfA, rfA), zA["solid"], QA.astype(np.float32), OA.grid_arrays(xfB, rfB), fl["T0"], fl["P0"])
    say(typ, "transfer onto", QB.shape, "cells")
    assert np.array_equal(QB.view(np.uint32), o.astype(np.float32).view(np.uint32))
    # the target steps on from the carried state
    B.step(10)
    assert np.isfinite(B.read()[0]).all()
    A.dispose(); B.dispose()


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
@pytest.mark.parametrize("case,tag", [("axi_tj_xcoarse", "it500"), ("axi_lbtf_xcoarse", "it100")])
def test_display_fields_and_stations_from_node(case, tag, tmp_path):
    """AxiHIP.derived and the station averages of js/cfd_post.js (over
    AxiHIP.sections) through the Node addon, against the oracle on the same
    FP32 states, to the bars above."""
    g, z, Q0, comps, xf, rf = load(case)
    z["mat"] = g.arr("z_mat")
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "axi_display_dump.js"), g.dir, tag, str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    grid = OA.grid_arrays(xf, rf)
    W, H, N = grid["W"], grid["H"], grid["W"] * grid["H"]
    Q = np.stack([g.arr(f"{tag}_Q{q}") for q in range(5)]).astype(np.float32)
    for disp in (0, 1):
        f = np.fromfile(tmp_path / f"der{disp}.f32", np.float32).reshape(13, N)
        reg = np.fromfile(tmp_path / f"reg{disp}.u8", np.uint8)
        mat = np.fromfile(tmp_path / f"mat{disp}.u8", np.uint8)
        fo, rego, mato = OA.derived(grid, z, Q, g.s["fl"]["T0"], g.s["fl"]["P0"], disp)
        assert np.array_equal(reg, rego) and np.array_equal(mat, mato)
        for c, key in enumerate(OA.DERIVED_KEYS):
            a, b = f[c], fo[c]
            if key in ("Pt", "ds"):
                fin = np.isfinite(b)
                assert np.array_equal(fin, np.isfinite(a)), key
                assert (np.abs(a[fin].astype(np.float64) - b[fin]) <= np.spacing(np.abs(b[fin])).astype(np.float64)).all(), (disp, key)
            else:
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (disp, key)
    # stations after one step, from the stepped state and mass flux the script read back
    Q1 = np.fromfile(tmp_path / "q1.f32", np.float32).reshape(5, N)
    mfx = np.fromfile(tmp_path / "mfx.f32", np.float32)
    st = json.load(open(tmp_path / "stations.json"))
    order = g.s["ctl"]["order"]
    typ = STATION_TYPES[case]
    assert [s["id"] for s in st] == [sid for sid, _, _ in STATIONS[typ]]
    for s, (sid, x, sn) in zip(st, STATIONS[typ]):
        col = col_at(grid["xc"], x)
        o = dict(zip(OA.SECTION_SUMS, OA.section_sums(grid, z, Q1, mfx, col, rows_of(z, W, H, col, order.index(sn)))))
        assert s["mdot"] == o["mdot"]
        if o["w"]:
            assert s["Tt"] == o["Tt"] / o["w"] and rel(s["Pt"], o["Pt"] / o["w"]) <= 2e-15, (sid, s, o)
        else:
            assert s["Tt"] is None and s["Pt"] is None       # 0 / 0 -> NaN, JSON null
    say(case, "display fields and", len(st), "stations through Node")
