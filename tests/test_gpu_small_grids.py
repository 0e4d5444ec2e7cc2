"""HIP solver on grids of one cell to a few tiles (tests/synth_cases.py): every
stage-kernel shape and the batched form against the oracle, which is pinned to
the reference's golden vectors. The grids are narrower than a wavefront row,
shorter than a tile and a z chunk, and just past a tile edge, with cut,
solid, baffle and fixed-state cells and the y = 0 mirror in every case.

Bar: 1e-5 relative per variable after 10 iterations (the reference's GPU/CPU
convention, tools/gpucmp.mjs:19-21). The scale is floored at 1e-3 of the
freestream's largest component, because with a handful of fluid cells and no
sideslip rho v stays at round-off level. The batched form must equal the
plain path bit for bit.
"""
import sys

import numpy as np
import pytest

import oracle3d as O
from parity import rel_diff
from synth_cases import SIZES, synth_case

pytestmark = pytest.mark.gpu

TOL = 1e-5
ITERS = 10
SHAPES = [(8, 0), (4, 0), (8, 1), (4, 3), (8, -3), (0, 1), (32, 0), (32, 2)]   # ty 32: tiles of 32 x 8


def say(*a):
    print("[gpu-small]", *a, flush=True)
    sys.stdout.flush()


def oracle_state(grid, geo, F, fix):
    o = O.Oracle3D(grid, geo)
    o.set_free(F); o.set_fix(fix); o.set_opt(0.5, 2); o.reset(); o.step(ITERS)
    return o.read()


@pytest.mark.parametrize("n", SIZES, ids=lambda n: "x".join(map(str, n)))
def test_small_grid_every_shape_matches_oracle(n):
    import aero3d
    grid, geo, F, fix = synth_case(*n, seed=sum(n))
    ref = oracle_state(grid, geo, F, fix)
    floor = 1e-3 * max(abs(f) for f in F)
    s = aero3d.CfdHIP(grid, geo)
    s.opt["cfl"] = 0.5
    s.set_free(F); s.set_fix(fix)
    for ty, zc in SHAPES:
        say("grid", n, "shape", ty, zc)
        s.set_tile(ty, zc)
        s.reset(); s.step(ITERS)
        d = rel_diff(s.read(), ref, geo["t"], floor=floor)
        say("grid", n, "shape", ty, zc, d)
        assert max(d.values()) < TOL, (n, ty, zc, d)
    s.dispose()


@pytest.mark.parametrize("mach", [2.0, -2.0])
def test_supersonic_freestream_matches_oracle(mach):
    """Mach 2 along +x and along -x: most x faces take the HLLC's upwind-side
    (supersonic) branches, left and right respectively, on every shape."""
    import aero3d
    n = (65, 9, 5)
    grid, geo, F, fix = synth_case(*n, seed=sum(n), mach=mach)
    ref = oracle_state(grid, geo, F, fix)
    floor = 1e-3 * max(abs(f) for f in F)
    s = aero3d.CfdHIP(grid, geo)
    s.opt["cfl"] = 0.5
    s.set_free(F); s.set_fix(fix)
    for ty, zc in [(8, 0), (4, 2), (32, 0)]:
        say("mach", mach, "shape", ty, zc)
        s.set_tile(ty, zc)
        s.reset(); s.step(ITERS)
        d = rel_diff(s.read(), ref, geo["t"], floor=floor)
        say("mach", mach, "shape", ty, zc, d)
        assert max(d.values()) < TOL, (mach, ty, zc, d)
    s.dispose()


@pytest.mark.parametrize("ty", [8, 4])
def test_balanced_launch_priority_bit_identical(ty):
    """A balanced launch that fits one resident round sets each wave's priority
    from its chunk's progress; a launch of three times as many blocks (several
    rounds) keeps the default issue order. On a 174 x 77 x 40 grid with a body
    the second launch has more z chunks per tile than the first (up to one per
    plane), so both forms run. Priority changes only the order of issue: after 20
    iterations the states and the last stage's local dt are the same bits."""
    import aero3d
    grid, geo, F, fix = synth_case(174, 77, 40, seed=3)
    out = []
    for zc in (0, -3):
        say("balanced launch", ty, zc)
        s = aero3d.CfdHIP(grid, geo)
        s.opt["cfl"] = 0.5
        s.set_free(F); s.set_fix(fix); s.reset()
        s.set_tile(ty, zc)
        s.step(20)
        out.append(s.read(with_dt=True))
        s.dispose()
    assert np.array_equal(out[0], out[1])


def test_stage_modes_bit_identical():
    """The balanced kernels are compiled once per stage (stage 1, stage 2, the
    last stage 2 writing dt); the fixed-chunk tile kernel decides the stage at
    run time. The same arithmetic either way: after 20 iterations the states
    and the local dt of both agree bit for bit, for 64 x 8, 64 x 4 and 32 x 8
    tiles (which also agree with each other)."""
    import aero3d
    grid, geo, F, fix = synth_case(174, 77, 40, seed=5)
    out = {}
    for ty, zc in ((8, 0), (8, 8), (4, 0), (4, 3), (32, 0), (32, 5)):
        say("stage modes", ty, zc)
        s = aero3d.CfdHIP(grid, geo)
        s.opt["cfl"] = 0.5
        s.set_free(F); s.set_fix(fix); s.reset()
        s.set_tile(ty, zc)
        s.step(20)
        out[(ty, zc)] = s.read(with_dt=True)
        s.dispose()
    ref = out[(8, 0)]
    for key, A in out.items():
        assert np.array_equal(A, ref), key


def test_small_grids_batched_equal_plain():
    import aero3d
    cs = [(n, *synth_case(*n, seed=sum(n))) for n in SIZES]
    say("batch of", len(cs), "small grids")
    b = aero3d.Batch(arena_bytes=1 << 28, max_slots=32)
    slots = {}
    for n, grid, geo, F, fix in cs:
        slots[n] = b.add(b.add_geometry(grid, geo), F, fix, 0.5)
    b.step(ITERS)
    for n, grid, geo, F, fix in cs:
        s = aero3d.CfdHIP(grid, geo)
        s.opt["cfl"] = 0.5
        s.set_free(F); s.set_fix(fix); s.reset(); s.step(ITERS)
        Qp = s.read()
        s.dispose()
        Qb = b.read(slots[n])
        assert b.iter(slots[n]) == ITERS
        assert np.array_equal(Qb.view(np.uint32), Qp.view(np.uint32)), f"{n}: batch differs from the plain path"
    b.dispose()


@pytest.mark.parametrize("ty,zc", SHAPES)
def test_freestream_and_rest_are_preserved(ty, zc):
    """The invariants of tests/test_oracle_invariants.py on the HIP path, every
    kernel shape, 200 iterations: a body-free grid stays at the freestream, and
    gas at rest around the cut-cell body stays at rest (bars 1e-6: FP32 with
    hardware reciprocals and fused multiply-adds)."""
    import aero3d
    for kw in ({"body": False}, {"quiescent": True}):
        grid, geo, F, fix = synth_case(65, 9, 5, seed=1, **kw)
        say("invariant", kw, "shape", ty, zc)
        s = aero3d.CfdHIP(grid, geo)
        s.opt["cfl"] = 0.5
        s.set_tile(ty, zc)
        s.set_free(F); s.set_fix(fix); s.reset(); s.step(200)
        Q = s.read().astype(np.float64)
        s.dispose()
        if "body" in kw:
            F32 = np.array(F, np.float32).astype(np.float64)
            dev = max(np.abs(Q[c] - F32[c]).max() for c in range(5)) / max(abs(f) for f in F)
            say("freestream deviation", dev)
            assert dev < 1e-6, (ty, zc, dev)
        else:
            fl = geo["t"] == 0
            u = max(np.abs(Q[c][fl]).max() for c in (1, 2, 3))
            say("rest: max |rho u|", u)
            assert u < 1e-6, (ty, zc, u)
            assert np.abs(Q[0][fl] - 1.0).max() < 1e-6

