"""The JS host layer (3d_aero_lab_amd/js) as a drop-in for the reference's entry points.

CPU: grid faces, cut-cell masks (t, phi, ap0..ap3) and force lists built by
the JS host are byte-identical to the reference's cfdGrid / cfdGeom /
cfdForceList outputs (tests/golden), the worker-thread build (GeomPool) gives
the same masks whatever the slab split, and the Node-API addon loads and throws
a JS Error when no device is usable (the reference's constructor contract,
src/air_cfd.js:371-377, 506-512).
GPU: the reference driver flow (tools/test_cfd3.js) through CfdHIP reproduces
the README row (fighter, 30k cells, M0.6, alpha 3, delta 3, 200 iterations).
"""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import geometry
from golden_io import Golden, sha256

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")


@pytest.mark.parametrize("case,spec", [("sphere", "sphere"), ("fighter30k", "fighter:30000:3"),
                                       ("interceptor", "interceptor:45000:-4"), ("transport60k", "transport:60000"),
                                       ("transport1m", "transport:1000000")])
def test_js_masks_byte_identical(case, spec):
    g = Golden(case)
    grid, geo, fl, meta = geometry.build_case(spec)
    arrs = {"fx": grid["fx"], "fy": grid["fy"], "fz": grid["fz"], "t": geo["t"], "phi": geo["phi"],
            **{f"ap{d}": geo["ap"][d] for d in range(4)}, **{f"fl_{k}": fl[k] for k in ("q", "cx", "cz", "x", "z")}}
    for k, a in arrs.items():
        assert sha256(a) == g.sha(k), f"{case}: {k} differs from the reference"
    assert meta["nSolid"] == g.s["nSolid"] and meta["nCut"] == g.s["nCut"]


def test_geom_pool_masks_byte_identical():
    """The worker-thread build (3 workers + the caller, slabs of 1, 5 and the default
    plane count) gives the reference's masks for every aircraft case."""
    cases = [("fighter30k", "fighter:30000:3"), ("interceptor", "interceptor:45000:-4"),
             ("transport60k", "airliner:60000:0:transport"), ("transport1m", "airliner:1000000:0:transport")]
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "geom_pool_check.js"), "3", *[s for _, s in cases]],
                       cwd=ROOT, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert len(res) == 3 * len(cases)
    for (case, spec), k in zip(cases, range(0, len(res), 3)):
        g = Golden(case)
        for b in res[k:k + 3]:
            assert b["spec"] == spec
            for name in ("t", "phi", "ap0", "ap1", "ap2", "ap3"):
                assert b[name] == g.sha(name), f"{case} slab {b['slab']}: {name} differs from the reference"
            assert b["nSolid"] == g.s["nSolid"] and b["nCut"] == g.s["nCut"]


def test_addon_throws_without_device():
    js = ("const {CfdHIP}=require('./3d_aero_lab_amd/js/hip_solver.js');const K=require('./3d_aero_lab_amd/js/cases.js');"
          "const C=K.sphereCase();try{new CfdHIP(C.g,C.geo).dispose();process.exit(3)}catch(e){console.log(e.message)}")
    r = subprocess.run(["node", "-e", js], cwd=ROOT, capture_output=True, text=True, timeout=120)
    if r.returncode == 3:
        pytest.skip("a device is present")
    assert r.returncode == 0 and "no HIP device" in r.stdout


def test_addon_exports():
    js = "const a=require('./3d_aero_lab_amd/lib/aero3d.node');console.log(JSON.stringify(Object.keys(a).sort()))"
    r = subprocess.run(["node", "-e", js], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    keys = json.loads(r.stdout)
    for k in ("create", "setFree", "setFix", "setOpt", "reset", "step", "read", "write", "dispose", "forces", "trefftz", "residual"):
        assert k in keys


def test_addon_rejects_foreign_handles():
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "addon_guard_check.js"), "cpu"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert all(v.startswith("TypeError") and "handle" in v for v in res.values()), res


@pytest.mark.gpu
def test_addon_argument_guards():
    """Short arrays, a handle of the other solver kind, use after dispose and a
    second dispose are JS errors, never a read or write past a buffer."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "addon_guard_check.js"), "gpu"], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print("[addon-guards]", res, flush=True)
    for k in ("create_short_phi", "read_short_output", "write_short_input", "force_list_short_column"):
        assert res[k].startswith("RangeError"), (k, res[k])
    assert res["axi_call_with_3d_handle"].startswith("TypeError"), res
    assert res["read_full_size"] == "no error", res
    assert "disposed" in res["step_after_dispose"] and "disposed" in res["dispose_twice"], res


@pytest.mark.gpu
def test_js_dropin_fighter_readme_row():
    g = Golden("fighter30k")
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "cfd3_hip_run.js"), "fighter", "30000", "200", "0.6", "3", "3", "0.7"],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    print("[js-dropin]", res, flush=True)
    assert res["N"] == g.s["N"]
    for h, ref in zip(res["hist"], g.s["hist"]):
        assert h["it"] == ref["it"]
        for k in ("CL", "CD"):
            assert abs(h["dev"][k] - ref[k]) / abs(ref[k]) < 1e-4
    assert round(res["hist"][-1]["dev"]["CL"], 4) == 0.1927 and round(res["hist"][-1]["dev"]["CD"], 5) == 0.04406
    assert abs(res["cdi"] - g.s["CDi"]) / g.s["CDi"] < 1e-4


@pytest.mark.gpu
def test_js_server_and_batch_from_node():
    """The served and batched forms driven from Node through the addon's async
    calls (serverTickAsync, batchStepAsync, stepAsync): every served request
    equals a plain CfdHIP solve read back by CfdMonitor under the same rule,
    bit for bit; a handle with an async call in flight refuses other calls."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "serve_node_check.js")], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    print("[js-serve]", out, flush=True)
    assert out["ok"], out
