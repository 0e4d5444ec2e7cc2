"""Device display sampling (aero3d_probe) against the reference's own
cfdStencil / cfdSample / cfdState / cfdVelAt (src/air_cfd.js:453-490,
637-646), run by tests/golden/make_golden.js on the reference's fighter state
after 200 iterations at 3000 seeded points: full, partial and nearest-fluid
stencils, points below the symmetry plane and outside the box of centres.

Bars: stencil cells, FP32 weights, found flags, primitives and particle
velocities bit for bit (the same FP64 operations in the same order); the
dimensional state bit for bit where it is a product or quotient of those
(P, T, rho, Cp, Vx, Vy, Vz) and within 1e-14 relative where it goes through
the square root of a sum of squares or a power (V, M, Tt, Pt, Taw: the
reference uses Math.hypot and Math.pow, whose last bit is the JS engine's).
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from golden_io import Golden

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXACT_KEYS = ("P", "T", "rho", "Cp", "Vx", "Vy", "Vz")
ULP_KEYS = ("V", "M", "Tt", "Pt", "Taw")


def say(*a):
    print("[gpu-probe]", *a, flush=True)
    sys.stdout.flush()


def fixed_states(t, Q):
    """The fixed states of cell types 3, 4, 5 as the state holds them (write()
    re-applies the fixed states, so they must be the run's own)."""
    out = []
    for ty in (3, 4, 5):
        k = np.flatnonzero(t == ty)
        out.append(Q[:, k[0]].astype(np.float64) if k.size else Q[:, 0].astype(np.float64))
    return out


def test_probe_equals_reference_sampling():
    import aero3d
    g, f = Golden("sampling"), Golden("fighter30k")
    assert np.array_equal(g.arr("t"), f.arr("t")) and np.array_equal(g.arr("fx"), f.arr("fx"))
    s = aero3d.CfdHIP(f.grid(), f.geo())
    Q = g.state("it200")
    s.set_fix(fixed_states(g.arr("t"), Q))
    s.write(Q)
    assert np.array_equal(s.read().view(np.uint32), Q.view(np.uint32))
    pts = g.arr("pts").reshape(-1, 3)
    n = pts.shape[0]
    say("probing", n, "points on the fighter state at 200 iterations")
    r = s.probe(pts, g.s["fs"], g.s["M"])
    s.dispose()
    ok = g.arr("ok").astype(bool)
    assert np.array_equal(r["ok"].astype(bool), ok)
    assert np.array_equal(r["ok"], g.arr("got")), "a found stencil always samples"
    assert np.array_equal(r["idx"].reshape(-1), g.arr("idx"))
    assert np.array_equal(r["w"].reshape(-1).view(np.uint32), g.arr("w").view(np.uint32))
    W = g.arr("W").reshape(-1, 5)
    bad = np.flatnonzero((r["W"][ok].view(np.uint64) != W[ok].view(np.uint64)).any(axis=1))
    assert bad.size == 0, (bad.size, bad[:5], r["W"][ok][bad[:3]], W[ok][bad[:3]])
    assert np.isnan(r["W"][~ok]).all()
    st = g.arr("state").reshape(-1, 12)
    keys = aero3d.PROBE_STATE_KEYS
    assert keys == g.s["keys"]
    worst = {}
    for c, k in enumerate(keys):
        a, b = r["state"][ok, c], st[ok, c]
        if k in EXACT_KEYS:
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), k
        else:
            d = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
            worst[k] = float(d.max())
            assert d.max() <= 1e-14, (k, d.max())
    vok = g.arr("vok").astype(bool)
    assert np.array_equal(r["vok"].astype(bool), vok)
    vel = g.arr("vel").reshape(-1, 3)
    assert np.array_equal(r["vel"][vok].view(np.uint64), vel[vok].view(np.uint64))
    assert np.isnan(r["vel"][~vok]).all()
    say("found", int(ok.sum()), "velocities", int(vok.sum()), "largest relative difference of the sqrt / pow fields", worst)
    # the fallback and mirrored branches are exercised
    assert (~vok & ok).any() and (pts[:, 1] < 0).any()


def test_probe_without_state_and_empty():
    import aero3d
    f = Golden("fighter30k")
    s = aero3d.CfdHIP(f.grid(), f.geo())
    F = [1, 0.5, 0, 0, 1 / (1.4 * 0.4) + 0.125]
    s.set_free(F); s.set_fix([F, F, F]); s.reset()
    r = s.probe(np.zeros((0, 3)))
    assert r["ok"].size == 0
    # a uniform state samples to the freestream primitives wherever a stencil is found
    far = np.array([[-200.0, 30.0, 40.0], [300.0, -50.0, -80.0]])
    r = s.probe(far, None, 0.5)
    s.dispose()
    assert r["state"] is None and r["ok"].all()
    np.testing.assert_allclose(r["W"][:, :4], [[1, 0.5, 0, 0]] * 2, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(r["W"][:, 4], 1 / 1.4, rtol=1e-6)


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_probe_from_node_equals_reference_sampling():
    """The same comparison through the Node addon (CfdHIP.probe of js/hip_solver.js)."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "sampling_check.js"), os.path.join(ROOT, "tests", "golden", "sampling"),
                        os.path.join(ROOT, "tests", "golden", "fighter30k")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout)
    say("node", out)
    assert out.pop("n") == 3000
    worst = out.pop("worst")
    assert all(out.values()), out
    assert max(worst.values()) <= 1e-14, worst
