"""Batched and served forms against the plain path and the oracle.

The batched kernel runs the same per-cell arithmetic as the plain kernel, so a
case gives bit-identical states whether it runs on the plain path, alone in a
batch, or mixed with other cases; and a served request gives the same result
(iterations, residuals, coefficients) alone and in a mix.
"""
import sys

import numpy as np
import pytest

from golden_io import Golden
from parity import rel_diff

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-batch]", *a, flush=True)
    sys.stdout.flush()


def cases():
    out = []
    for name in ("sphere", "transport60k", "interceptor"):
        g = Golden(name)
        F = g.s["F"]
        out.append((name, g, F, g.s.get("fix", [F, F, F]), g.s["cfl"]))
    return out


def plain(g, F, fix, cfl, n):
    import aero3d
    s = aero3d.CfdHIP(g.grid(), g.geo())
    s.opt["cfl"] = cfl
    s.set_free(F); s.set_fix(fix); s.reset(); s.step(n)
    Q = s.read()
    s.dispose()
    return Q


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("group_cells", [2_000_000, 40_000])
def test_batch_equals_plain_and_oracle(group_cells, persistent):
    """Both launch forms (one block per work item, resident blocks walking the
    items), one group and one group per case: bit-identical to the plain path,
    before and after a case leaves (which changes the item count)."""
    import aero3d
    b = aero3d.Batch(arena_bytes=1 << 30, max_slots=16)
    b.set_group_cells(group_cells)        # 40k: every case in its own group
    b.set_persistent(persistent)
    say("batch", group_cells, "persistent", persistent)
    cs = cases()
    gids = {name: b.add_geometry(g.grid(), g.geo()) for name, g, *_ in cs}
    slots = {}
    for name, g, F, fix, cfl in cs:
        slots[name] = b.add(gids[name], F, fix, cfl)
    # a second case on the shared sphere geometry, other Mach number
    Fs = list(cs[0][2]); Fs[1] = 0.3; Fs[4] = 1 / (1.4 * 0.4) + 0.045
    s2 = b.add(gids["sphere"], Fs, [Fs] * 3, 0.5)
    say("footprint", b.footprint())
    b.step(7)       # two graph replays + one direct iteration
    b.step(3)
    for name, g, F, fix, cfl in cs:
        Q = b.read(slots[name])
        assert b.iter(slots[name]) == 10
        Qp = plain(g, F, fix, cfl, 10)
        assert np.array_equal(Q.view(np.uint32), Qp.view(np.uint32)), f"{name}: batch differs from plain path"
        if name == "sphere":
            d = rel_diff(Q, g.state("it10"), g.geo()["t"])
            say(name, d)
            assert max(d.values()) < 1e-5
    Qs2 = b.read(s2)
    Qp2 = plain(cs[0][1], Fs, [Fs] * 3, 0.5, 10)
    assert np.array_equal(Qs2.view(np.uint32), Qp2.view(np.uint32))
    # remove a case mid-run: the others continue unchanged
    b.remove(slots["interceptor"])
    b.step(90)
    Q = b.read(slots["transport60k"])
    g = cs[1][1]
    d = rel_diff(Q, g.state("it100"), g.geo()["t"])
    say("transport60k it100 in batch", d)
    assert max(d.values()) < 1e-5
    assert np.array_equal(Q.view(np.uint32), plain(g, cs[1][2], cs[1][3], cs[1][4], 100).view(np.uint32))
    b.dispose()


def test_alone_equals_mixed():
    import aero3d
    cs = cases()
    name, g, F, fix, cfl = cs[1]
    alone = aero3d.Batch(arena_bytes=1 << 29, max_slots=4)
    ga = alone.add_geometry(g.grid(), g.geo())
    sa = alone.add(ga, F, fix, cfl)
    alone.step(25)
    Qa = alone.read(sa)
    ra = alone.residuals([sa])
    alone.dispose()
    mix = aero3d.Batch(arena_bytes=1 << 30, max_slots=8)
    gids = [mix.add_geometry(c[1].grid(), c[1].geo()) for c in cs]
    for k, c in enumerate(cs):
        mix.add(gids[k], c[2], c[3], c[4])
    mix.step(5)                                   # others start earlier
    sm = mix.add(gids[1], F, fix, cfl)
    mix.step(25)
    assert np.array_equal(mix.read(sm).view(np.uint32), Qa.view(np.uint32))
    mix.dispose()
    del ra


def test_server_alone_equals_mixed_and_admission_by_footprint():
    import aero3d
    cs = cases()
    g = cs[1][1]
    r = g.s["ref"]
    fref = (g.s["alpha"], g.s["M"], r["S"], r["xcg"], 0.0, r["c"])

    def serve(reqs, arena):
        srv = aero3d.Server(arena_bytes=arena, max_slots=8, check_every=50)
        gid = {c[0]: srv.batch.add_geometry(c[1].grid(), c[1].geo(), c[1].force_list()) for c in cs}
        for tag, (name, F, fix, cfl, it) in enumerate(reqs):
            ref = fref if name == "transport60k" else (0.0, 0.4, 1.0, 0.0, 0.0, 1.0)
            srv.submit(gid[name], F, fix, cfl, 2, ref, 1e-4, it, srv.clock(), tag)
        out, ticks = [], 0
        while len(out) < len(reqs):
            a, p = srv.tick()
            ticks += 1
            out += srv.poll()
            assert ticks < 1000
        srv.dispose()
        return {o["tag"]: o for o in out}

    F, fix, cfl = cs[1][2], cs[1][3], cs[1][4]
    alone = serve([("transport60k", F, fix, cfl, 150)], 1 << 30)
    mixed_reqs = [("sphere", cs[0][2], cs[0][3], 0.5, 100), ("interceptor", cs[2][2], cs[2][3], 0.5, 200),
                  ("transport60k", F, fix, cfl, 150), ("sphere", cs[0][2], cs[0][3], 0.5, 300)]
    # arena too small to hold every case at once: later requests wait for memory
    import aero3d as A3
    b = A3.Batch(arena_bytes=1 << 30, max_slots=2)
    geo_bytes = 0
    need = []
    for c in cs:
        gi = b.add_geometry(c[1].grid(), c[1].geo(), c[1].force_list())
        need.append(b.case_bytes(gi))
    geo_bytes = b.footprint()[0]
    b.dispose()
    arena = geo_bytes + max(need) + min(need) + 4096
    mixed = serve(mixed_reqs, arena)
    say("alone", alone[0])
    say("mixed", mixed[2])
    for k in ("iters", "converged", "res0", "res", "CL", "CD", "Cm"):
        assert alone[0][k] == mixed[2][k], k
    assert mixed[0]["iters"] == 100 and mixed[3]["iters"] == 300
    # requests waited for memory: at least one was admitted after another finished
    admits = sorted(o["t_admit"] for o in mixed.values())
    dones = sorted(o["t_done"] for o in mixed.values())
    assert admits[-1] >= dones[0]


def test_server_requests_that_cannot_fit_do_not_stall():
    """A request larger than the arena beside the registered geometries is
    refused at submit; one that a later geometry leaves without room is
    returned as failed (iters 0, NaN coefficients) instead of waiting forever,
    and the queue behind it is served."""
    import aero3d
    g = Golden("sphere")
    F = g.s["F"]
    probe = aero3d.Batch(arena_bytes=1 << 28, max_slots=2)
    gi = probe.add_geometry(g.grid(), g.geo())
    need, geo_bytes = probe.case_bytes(gi), probe.footprint()[0]
    probe.dispose()
    say("sphere case bytes", need, "geometry bytes", geo_bytes)
    ref = (0.0, 0.4, 1.0, 0.0, 0.0, 1.0)
    # arena smaller than geometry + one case: refused at submit
    srv = aero3d.Server(arena_bytes=geo_bytes + need // 2, max_slots=4, check_every=10)
    gid = srv.batch.add_geometry(g.grid(), g.geo())
    with pytest.raises(aero3d.Aero3DError, match="more than the arena"):
        srv.submit(gid, F, [F] * 3, 0.5, 2, ref, 1e-4, 20, srv.clock(), 7)
    assert srv.tick() == (0, 0)
    srv.dispose()
    # room for one case; a second geometry registered after the submit takes it
    srv = aero3d.Server(arena_bytes=geo_bytes + need + 8192, max_slots=4, check_every=10)
    gid = srv.batch.add_geometry(g.grid(), g.geo())
    srv.submit(gid, F, [F] * 3, 0.5, 2, ref, 1e-4, 20, srv.clock(), 1)
    gid2 = srv.batch.add_geometry(g.grid(), g.geo())
    a, p = srv.tick()
    out = srv.poll()
    say("stranded request", out)
    assert (a, p) == (0, 0) and len(out) == 1 and out[0]["tag"] == 1 and out[0]["iters"] == 0
    assert np.isnan(out[0]["CL"]) and out[0]["converged"] == 0
    with pytest.raises(aero3d.Aero3DError, match="more than the arena"):
        srv.submit(gid2, F, [F] * 3, 0.5, 2, ref, 1e-4, 20, srv.clock(), 2)
    srv.dispose()


def test_zchunk_policies_bit_identical():
    """The z chunk only changes which block computes which planes: fixed chunks
    of 1, 3 and 8 planes and the chunk chosen from the load (0, the default)
    give the same bits as the plain path; a negative chunk is an argument
    error."""
    import aero3d
    cs = cases()
    ref = {name: plain(g, F, fix, cfl, 12) for name, g, F, fix, cfl in cs}
    for zc in (0, 1, 3, 8):
        b = aero3d.Batch(arena_bytes=1 << 30, max_slots=8)
        gids = {name: b.add_geometry(g.grid(), g.geo()) for name, g, *_ in cs}
        slots = {name: b.add(gids[name], F, fix, cfl) for name, g, F, fix, cfl in cs}
        b.set_zchunk(zc)
        b.step(7); b.step(5)
        for name in slots:
            Q = b.read(slots[name])
            assert np.array_equal(Q.view(np.uint32), ref[name].view(np.uint32)), (zc, name)
        if zc == 0:
            with pytest.raises(aero3d.Aero3DError):
                b.set_zchunk(-1)
        say("zchunk", zc, "bit-identical to the plain path")
        b.dispose()
