"""Convergence rules of a monitored solve against the reference controller.

The reference's own readback controller (CFD3.readback, src/air_cfd.js:550-578)
was run to convergence on the sphere and the fighter by
tests/golden/make_golden.js; its readback history (residual, CL, the conv
flag at every readback) is the fixture. Each implementation of the rule must
raise the flag at the same readback: the device rule of the served form
(csrc/converge.hpp, compiled here for the host), the Python mirror
(aero3d.ConvRule) and the JS mirror of the solver object (js/hip_solver.js).
Edge cases the reference's runs do not reach are pinned by construction:
a NaN residual or lift coefficient never converges, the cap ends a solve.
"""
import ctypes as C
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from golden_io import Golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "3d_aero_lab_amd", "csrc")
BUILD = os.path.join(ROOT, "tests", "cpu_emul", "_build")


@pytest.fixture(scope="module")
def crule():
    os.makedirs(BUILD, exist_ok=True)
    so = os.path.join(BUILD, "libconvrule.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-I", CSRC, "-o", so,
                    os.path.join(ROOT, "tests", "cpu_emul", "conv_rule.cpp")], check=True)
    L = C.CDLL(so)
    d, i = np.ctypeslib.ndpointer(np.float64), np.ctypeslib.ndpointer(np.int32)
    L.conv_run.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, d, d, i, i, i, i, d]

    def run(rule, res_drop, cap, res, cl, it):
        n = len(res)
        out = [np.zeros(n, np.int32) for _ in range(3)] + [np.zeros(n)]
        L.conv_run(rule, res_drop, cap, n, np.asarray(res, np.float64), np.asarray(cl, np.float64), np.asarray(it, np.int32), *out)
        return out
    return run


def reads(case):
    r = Golden(case).s["reads"]
    res = [float("nan") if x["res"] is None else x["res"] for x in r]
    return r, res, [x["CL"] for x in r], [x["it"] for x in r]


@pytest.mark.parametrize("case", ["readback_sphere", "readback_fighter30k", "readback_transport60k"])
def test_device_rule_matches_reference_controller(crule, case):
    r, res, cl, it = reads(case)
    conv, div, done, res0 = crule(1, 0.0, 30000, res, cl, it)
    assert [bool(c) for c in conv] == [x["conv"] for x in r]
    assert [bool(d) for d in done] == [x["conv"] for x in r]      # no cap or divergence in these runs
    assert not div.any()
    ref0 = [float("nan") if x["res0"] is None else x["res0"] for x in r]
    assert all((a == b) or (math.isnan(a) and math.isnan(b)) for a, b in zip(res0, ref0))


@pytest.mark.parametrize("case", ["readback_sphere", "readback_fighter30k", "readback_transport60k"])
def test_python_rule_matches_reference_controller(case):
    import aero3d
    r, res, cl, it = reads(case)
    rule = aero3d.ConvRule(aero3d.RULE_REFERENCE, cap=30000)
    flags = [rule.update(a, b, c) and rule.conv for a, b, c in zip(res, cl, it)]
    assert flags == [x["conv"] for x in r]


def test_converged_runs_end_where_the_reference_does():
    """The two long runs converge (the flag is raised once, at the last readback)."""
    for case, it_end in (("readback_sphere", 1200), ("readback_fighter30k", 1250)):
        r = Golden(case).s["reads"]
        assert r[-1]["conv"] and not any(x["conv"] for x in r[:-1]) and r[-1]["it"] == it_end


def test_rule_edge_cases(crule):
    import aero3d
    seq_it = [50 * k for k in range(12)]
    # a residual that falls geometrically: converged at the first readback past six with res/res0 < 2e-3
    res = [float("nan")] + [1e-3 * 0.25 ** k for k in range(11)]
    cl = [0.3] * 12
    for kind, fn in (("c", None), ("py", None)):
        if kind == "c":
            conv, div, done, _ = crule(1, 0.0, 30000, res, cl, seq_it)
            conv = list(map(bool, conv)); div = list(map(bool, div)); done = list(map(bool, done))
        else:
            R = aero3d.ConvRule(aero3d.RULE_REFERENCE, cap=30000)
            conv, div, done = [], [], []
            for a, b, c in zip(res, cl, seq_it):
                R.update(a, b, c); conv.append(R.conv); div.append(R.diverged); done.append(R.done)
        first = conv.index(True)
        assert first == 6 and res[first] / res[1] < 2e-3 and res[first - 1] / res[1] >= 2e-3
        assert done == conv and not any(div)
    # a NaN lift coefficient in the last four readbacks (readbacks 7..10): no convergence there
    cl2 = list(cl); cl2[7] = float("nan")
    conv, _, _, _ = crule(1, 0.0, 30000, res, cl2, seq_it)
    assert list(conv) == [0] * 6 + [1] + [0] * 4 + [1]
    # a lift coefficient spread of 2e-3 or more over the last four: no convergence
    cl3 = [0.3 + (2e-3 if k % 2 else 0) for k in range(12)]
    conv, _, _, _ = crule(1, 0.0, 30000, res, cl3, seq_it)
    assert not conv.any()
    # a residual that turns NaN after the first readback: diverged and done, for both rules
    res4 = [float("nan"), 1e-3, 5e-4, float("nan")]
    for rule in (0, 1):
        conv, div, done, _ = crule(rule, 1e-4, 30000, res4, [0.1] * 4, [0, 50, 100, 150])
        assert list(div) == [0, 0, 0, 1] and list(done) == [0, 0, 0, 1] and not conv.any()
    # caps: the reference rule ends past the cap, the residual-drop rule at it
    conv, div, done, _ = crule(1, 0.0, 100, [float("nan"), 1e-3, 1e-3, 1e-3], [0.1] * 4, [0, 50, 100, 150])
    assert list(done) == [0, 0, 0, 1]
    conv, div, done, _ = crule(0, 1e-4, 100, [float("nan"), 1e-3, 1e-3], [0.1] * 3, [0, 50, 100])
    assert list(done) == [0, 0, 1] and not conv.any()
    conv, div, done, _ = crule(0, 1e-1, 1000, [float("nan"), 1e-3, 1e-4], [0.1] * 3, [0, 50, 100])
    assert list(conv) == [0, 0, 1] and list(done) == [0, 0, 1]


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_js_rule_matches_reference_controller():
    out = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "conv_rule_check.js")], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout)
    for case in ("readback_sphere", "readback_fighter30k", "readback_transport60k"):
        assert got[case] == [x["conv"] for x in Golden(case).s["reads"]], case
    assert got["edge"]["nan_res_diverged"] and got["edge"]["nan_cl_no_conv"] and got["edge"]["cap"]


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_axi_divergence_restart():
    """The axisymmetric loop's divergence handling (src/app_cfd.js:377-378,
    405-409) on a stand-in solver that plants a NaN density: the run restarts
    from the initial state with CFL - 0.2 (floored at 0.3) and first order;
    after maxRestarts restarts it gives up."""
    out = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "axi_divergence_check.js")], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout)
    once = r["once"]
    assert once["restarts"] == [{"at": 500, "cfl": 0.6, "order": 2}]
    assert once["made"] == [{"cfl": 0.6, "order": 2}, {"cfl": 0.6 - 0.2, "order": 1}]
    assert not once["diverged"] and once["iters"] == 3250          # the clean run converges past 3000 iterations
    a = r["always"]
    assert a["diverged"] and [x["cfl"] for x in a["made"]] == [0.6, 0.6 - 0.2, 0.3, 0.3]
    assert [x["order"] for x in a["made"]] == [2, 1, 1, 1] and len(a["restarts"]) == 4
