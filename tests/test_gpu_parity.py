"""HIP solver (libaero3d.so) against the CPU oracle and the reference's golden vectors.

Bars: cell flags and masks are inputs shared bit for bit with the reference;
conserved variables within 1e-5 relative per cell after N iterations (scale
convention of tools/gpucmp.mjs:19-21, FP32 device arithmetic against the
oracle's double arithmetic); CL / CD within 1e-4 relative.
"""
import sys

import numpy as np
import pytest

import oracle3d as O
from golden_io import Golden
from parity import rel, rel_diff

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-5
COEF_TOL = 1e-4


def say(*a):
    print("[gpu-parity]", *a, flush=True)
    sys.stdout.flush()


def make_hip(g, F, fix, cfl, order=2):
    import aero3d
    s = aero3d.CfdHIP(g.grid(), g.geo())
    s.opt["cfl"] = cfl
    s.opt["order"] = order
    s.set_free(F)
    s.set_fix(fix)
    s.reset()
    return s


@pytest.mark.parametrize("case,order,iters", [("sphere", 2, [1, 10, 200]), ("sphere_o1", 1, [10]),
                                              ("interceptor", 2, [20]), ("transport60k", 2, [2, 100])])
def test_state_matches_reference(case, order, iters):
    g = Golden(case)
    base = Golden("sphere") if case == "sphere_o1" else g
    if case == "sphere_o1":
        g.grid, g.geo = base.grid, base.geo
    F = base.s["F"]
    fix = g.s.get("fix", [F, F, F])
    cfl = base.s["cfl"]
    say(case, "order", order, "iters", iters)
    s = make_hip(g, F, fix, cfl, order)
    t = g.geo()["t"]
    done = 0
    for it in iters:
        s.step(it - done)
        done = it
        Q = s.read()
        d = rel_diff(Q, g.state(f"it{it}"), t)
        say(case, it, d)
        assert max(d.values()) < STATE_TOL, f"{case} it{it}: {d}"
    assert s.it == iters[-1]
    s.dispose()


def test_fighter_readme_coefficients():
    """README.md:98 row: CL 0.1927, CD 0.04406 after 200 iterations (CFL 0.7)."""
    g = Golden("fighter30k")
    F = g.s["F"]
    s = make_hip(g, F, [F, F, F], 0.7)
    s.set_force_list(g.force_list())
    r = g.s["ref"]
    done = 0
    for h in g.s["hist"]:
        s.step(h["it"] - done)
        done = h["it"]
        f = s.forces(g.s["alpha"], g.s["M"], r["S"], r["xcg"], 0, r["c"])
        say("fighter", h["it"], f, h)
        assert rel(f["CL"], h["CL"]) < COEF_TOL and rel(f["CD"], h["CD"]) < COEF_TOL
    Q = s.read()
    d = rel_diff(Q, g.state("it200"), g.geo()["t"])
    say("fighter state", d)
    assert max(d.values()) < STATE_TOL
    cdi = s.trefftz(g.s["xT"], g.s["alpha"], g.s["M"], r["S"])
    assert rel(cdi, g.s["CDi"]) < COEF_TOL


def test_transport_1m_matches_oracle():
    """The headline workload at full size: the 1.0M-cell transport (M 0.78,
    alpha 2, CFL 0.5, fixed nacelle states) after 200 iterations on the GPU
    against the oracle on the host cores: every cell within 1e-5 and CL / CD /
    Cm within 1e-4. The oracle takes about 8 s on 16 cores."""
    import math
    import os
    import aero3d
    import geometry
    say("transport 1M: grid and masks")
    grid, geo, fl, meta = geometry.build_case("transport:1000000")
    M, alpha = 0.78, 2.0
    a = math.radians(alpha)
    F = [1.0, M * math.cos(a), 0.0, M * math.sin(a), 1 / (1.4 * 0.4) + 0.5 * M * M]
    fix = meta["fix"]
    r = meta["ref"]
    fref = (alpha, M, r["S"], r["xcg"], 0.0, r["c"])
    s = aero3d.CfdHIP(grid, geo)
    s.opt["cfl"] = 0.5
    s.set_free(F); s.set_fix(fix); s.reset(); s.set_force_list(fl)
    s.step(200)
    Q = s.read()
    fd = s.forces(*fref)
    s.dispose()
    say("transport 1M: GPU done", fd, "; oracle 200 iterations")
    o = O.Oracle3D(grid, geo, threads=min(16, os.cpu_count() or 1))
    o.set_free(F); o.set_fix(fix); o.set_opt(0.5, 2); o.reset(); o.step(200)
    Qo = o.read()
    fo = O.forces(Qo, fl, *fref)
    d = rel_diff(Q, Qo, geo["t"])
    say("transport 1M state", d, "coefficients GPU", fd, "oracle", fo)
    assert max(d.values()) < STATE_TOL, d
    for k in ("CL", "CD", "Cm"):
        assert rel(fd[k], fo[k]) < COEF_TOL, (k, fd[k], fo[k])


@pytest.mark.timeout(900)
def test_transport_1m_coefficients_after_2000_iterations():
    """The 1.0M-cell transport carried to 2000 iterations, as the headline
    configuration is quoted (CL / CD / Cm of the reference's driver flow at
    2000 iterations): the GPU's coefficients against the oracle's within 1e-4.
    The oracle takes about 80 s on 16 cores; progress every 250 iterations."""
    import math
    import os
    import time
    import aero3d
    import geometry
    grid, geo, fl, meta = geometry.build_case("transport:1000000")
    M, alpha = 0.78, 2.0
    a = math.radians(alpha)
    F = [1.0, M * math.cos(a), 0.0, M * math.sin(a), 1 / (1.4 * 0.4) + 0.5 * M * M]
    fix = meta["fix"]
    r = meta["ref"]
    fref = (alpha, M, r["S"], r["xcg"], 0.0, r["c"])
    s = aero3d.CfdHIP(grid, geo)
    s.opt["cfl"] = 0.5
    s.set_free(F); s.set_fix(fix); s.reset(); s.set_force_list(fl)
    s.step(2000)
    Q = s.read()
    fd = s.forces(*fref)
    s.dispose()
    say("transport 1M, 2000 iterations: GPU", fd)
    o = O.Oracle3D(grid, geo, threads=min(16, os.cpu_count() or 1))
    o.set_free(F); o.set_fix(fix); o.set_opt(0.5, 2); o.reset()
    t0 = time.perf_counter()
    for done in range(250, 2001, 250):
        o.step(250)
        say("oracle at", done, "iterations,", round(time.perf_counter() - t0, 1), "s")
    Qo = o.read()
    fo = O.forces(Qo, fl, *fref)
    d = rel_diff(Q, Qo, geo["t"])
    say("transport 1M, 2000 iterations: state", d, "coefficients GPU", fd, "oracle", fo)
    for k in ("CL", "CD", "Cm"):
        assert rel(fd[k], fo[k]) < COEF_TOL, (k, fd[k], fo[k])


def test_device_reductions_match_oracle_on_same_state():
    """Forces, Trefftz and residual in FP64 on the device equal the oracle's
    sums over the identical FP32 state to summation-order rounding."""
    g = Golden("transport60k")
    F, fix = g.s["F"], g.s["fix"]
    s = make_hip(g, F, fix, 0.5)
    fl = g.force_list()
    s.set_force_list(fl)
    grid, geo = g.grid(), g.geo()
    assert np.isnan(s.residual())
    s.step(7)
    Q0 = s.read()
    s.step(5)
    Q = s.read()
    res = s.residual()
    r = g.s["ref"]
    f = s.forces(g.s["alpha"], g.s["M"], r["S"], r["xcg"], 0, r["c"])
    fo = O.forces(Q, fl, g.s["alpha"], g.s["M"], r["S"], r["xcg"], 0, r["c"])
    for k in ("CL", "CD", "Cm"):
        assert rel(f[k], fo[k]) < 1e-10, (k, f[k], fo[k])
    cdi = s.trefftz(g.s["xT"], g.s["alpha"], g.s["M"], r["S"])
    assert rel(cdi, O.trefftz(grid, geo["t"], Q, g.s["xT"], g.s["alpha"], g.s["M"], r["S"])) < 1e-10
    # residual snapshot was taken at it 0 (the reset state F); the next call measures it 0 -> 12
    Qf = np.stack([np.full(len(Q[0]), F[c], np.float32) for c in range(5)])
    assert rel(res, O.residual(geo["t"], Q[0], Qf[0], 12)) < 1e-10
    s.step(3)
    res2 = s.residual()
    assert rel(res2, O.residual(geo["t"], s.read()[0], Q[0], 3)) < 1e-10
    del Q0


def test_force_list_replaced():
    """A force list can be replaced on a live solver by a shorter and then a
    longer one (the device buffers are freed and allocated again); the device
    sums then cover exactly the new list, to summation-order rounding."""
    g = Golden("transport60k")
    F, fix = g.s["F"], g.s["fix"]
    s = make_hip(g, F, fix, 0.5)
    s.step(5)
    Q = s.read()
    fl = g.force_list()
    r = g.s["ref"]
    args = (g.s["alpha"], g.s["M"], r["S"], r["xcg"], 0, r["c"])
    half = {k: np.ascontiguousarray(v[: len(v) // 2]) for k, v in fl.items()}
    for lst in (fl, half, fl):
        s.set_force_list(lst)
        f = s.forces(*args)
        fo = O.forces(Q, lst, *args)
        say("force list", len(lst["q"]), f, fo)
        for k in ("CL", "CD", "Cm"):
            assert rel(f[k], fo[k]) < 1e-10, (len(lst["q"]), k, f[k], fo[k])
    s.dispose()


def test_write_read_roundtrip_and_step_zero():
    g = Golden("sphere")
    F = g.s["F"]
    s = make_hip(g, F, [F, F, F], 0.5)
    Q = g.state("it10")
    s.write(Q)
    s.step(0)
    assert np.array_equal(s.read(), Q)
    # continuing from the written state matches the oracle continuing from it
    o = O.Oracle3D(g.grid(), g.geo())
    o.set_free(F); o.set_fix([F, F, F]); o.set_opt(0.5, 2); o.write(Q)
    s.step(5); o.step(5)
    d = rel_diff(s.read(), o.read(), g.geo()["t"])
    assert max(d.values()) < STATE_TOL, d


def test_bad_option_is_an_error():
    import aero3d
    g = Golden("sphere")
    F = g.s["F"]
    s = make_hip(g, F, [F, F, F], 0.5)
    s.opt["order"] = 3
    with pytest.raises(aero3d.Aero3DError, match="order"):
        s.step(1)


@pytest.mark.parametrize("ty,zc", [(0, 1), (4, 3), (4, 75), (8, 1), (8, 8), (8, 80), (8, 0), (4, 0), (8, -3), (4, -8),
                                   (32, 0), (32, 3), (32, -2)])
def test_kernel_shapes_agree(ty, zc):
    """Every stage-kernel shape (per-cell kernel, tiles of 64 x 4 / 64 x 8 / 32 x 8 with
    short to whole-domain z chunks, balanced launches) reproduces the oracle."""
    g = Golden("transport60k")
    F, fix = g.s["F"], g.s["fix"]
    say("shape", ty, zc)
    s = make_hip(g, F, fix, 0.5)
    s.set_tile(ty, zc)
    s.step(2)
    d = rel_diff(s.read(), g.state("it2"), g.geo()["t"])
    say("shape", ty, zc, d)
    assert max(d.values()) < STATE_TOL, d


def test_autotune_keeps_state():
    g = Golden("sphere")
    F = g.s["F"]
    s = make_hip(g, F, [F, F, F], 0.5)
    s.step(10)
    Q = s.read(with_dt=True)
    ty, zc, ms = s.autotune(3)
    say("autotune sphere", ty, zc, ms)
    assert ty in (4, 8, 32) and zc >= -8 and ms > 0
    assert s.it == 10 and np.array_equal(s.read(with_dt=True), Q)      # state and the local dt restored
    s.step(190)
    d = rel_diff(s.read(), g.state("it200"), g.geo()["t"])
    assert max(d.values()) < STATE_TOL, d


def test_readback_matches_oracle_functions():
    """aero3d_readback (residual since the previous readback, coefficients,
    Trefftz CDi, Mach maximum in one call) against the oracle's functions on
    the GPU's own states: the sums to summation-order rounding, the Mach
    maximum (order-independent) bit for bit."""
    import aero3d
    g = Golden("readback_transport60k")
    b = Golden("transport60k")
    F, fix = g.s["F"], g.s["fix"]
    s = make_hip(b, F, fix, 0.5)
    fl = b.force_list()
    s.set_force_list(fl)
    grid, geo = b.grid(), b.geo()
    r = g.s["ref"]
    ref = (g.s["alpha"], g.s["M"], r["S"], r["xcg"], 0.0, r["c"])
    xT = g.s["reads"][0]["xT"]
    o0 = s.readback(*ref, xT)
    assert o0["it"] == 0 and np.isnan(o0["res"])
    s.step(30)
    Qa = s.read()
    o1 = s.readback(*ref, xT)
    s.step(20)
    Qb = s.read()
    o2 = s.readback(*ref, xT)
    assert o2["it"] == 50
    assert rel(o1["res"], O.residual(geo["t"], Qa[0], np.full(len(Qa[0]), F[0], np.float32), 30)) < 1e-12
    assert rel(o2["res"], O.residual(geo["t"], Qb[0], Qa[0], 20)) < 1e-12
    fo = O.forces(Qb, fl, *ref)
    for k in ("CL", "CD", "Cm"):
        assert rel(o2[k], fo[k]) < 1e-12, (k, o2[k], fo[k])
    assert rel(o2["CDi"], O.trefftz(grid, geo["t"], Qb, xT, g.s["alpha"], g.s["M"], r["S"])) < 1e-12
    assert o2["Mmax"] == O.mach_max(geo["t"], Qb)
    say("readback at 50", o2)
    # no force list, no Trefftz plane: NaN coefficients and CDi, the rest as before
    s2 = make_hip(b, F, fix, 0.5)
    s2.step(5)
    o = s2.readback(*ref)
    assert np.isnan(o["CL"]) and np.isnan(o["CDi"]) and o["Mmax"] == O.mach_max(geo["t"], s2.read())
    s.dispose(); s2.dispose()


@pytest.mark.parametrize("case,base", [("readback_sphere", "sphere"), ("readback_fighter30k", "fighter30k")])
def test_converges_like_the_reference_controller(case, base):
    """The reference's own readback controller (CFD3.readback, run to its
    convergence flag on CfdCPU) against the HIP solver read back every 50
    iterations under the same rule (aero3d.ConvRule, reference rule).
    At every common readback the coefficients, CDi and the Mach maximum agree
    within 1e-3 relative and the residual within 5 % (it is the difference of
    two states 50 iterations apart, so FP32 arithmetic shows in it first); the
    rule, whose residual ratio falls about 10 % per readback near 2e-3, fires
    within one readback of the reference's; the state at the reference's
    convergence iteration is within 1e-5 per cell. For the sphere this gives
    C_d of the reference's scheme at radius 4.8 cells on this grid, 0.4390
    (the README's 0.56, README.md:99, is quoted for a grid it does not state)."""
    import aero3d
    g, b = Golden(case), Golden(base)
    F = g.s["F"]
    s = make_hip(b, F, [F, F, F], g.s["cfl"])
    s.set_force_list(b.force_list())
    r = g.s["ref"]
    ref = (g.s["alpha"], g.s["M"], r["S"], r["xcg"], r.get("zcg") or 0.0, r["c"])
    rule = aero3d.ConvRule(aero3d.RULE_REFERENCE, cap=30000)
    hist = s.solve(ref, rule, 50, g.s["reads"][0]["xT"])
    last, want = hist[-1], g.s["reads"][-1]
    say(case, "HIP converged at", last["it"], "CL", last["CL"], "CD", last["CD"], "CDi", last["CDi"], "Mmax", last["Mmax"],
        "| reference at", want["it"], "CL", want["CL"], "CD", want["CD"], "CDi", want["CDi"], "Mmax", want["Mmax"])
    assert rule.conv and abs(last["it"] - want["it"]) <= 50
    mine = {h["it"]: h for h in hist}
    worst = {}
    for rd in g.s["reads"]:
        h = mine.get(rd["it"])
        if h is None or rd["res"] is None:
            continue
        keys = ("CD", "CDi", "Mmax") + (("CL", "Cm") if case != "readback_sphere" else ())
        for k in keys + ("res",):
            worst[k] = max(worst.get(k, 0.0), rel(h[k], rd[k]))
        say(case, rd["it"], "res", h["res"], rd["res"], "ratio", h["res"] / rule.res0, rd["res"] / rd["res0"], "CD", h["CD"], rd["CD"])
    say(case, "largest relative differences over the common readbacks", worst)
    for k, v in worst.items():
        assert v < (0.05 if k == "res" else 1e-3), (k, v)
    if case == "readback_sphere":
        assert max(abs(h["CL"]) for h in hist) < 1e-5
    s.reset()
    s.step(want["it"])
    d = rel_diff(s.read(), g.state("final"), b.geo()["t"])
    say(case, "state at the reference's convergence iteration", want["it"], d)
    assert max(d.values()) < STATE_TOL, d
    s.dispose()
