"""The served form (continuous batching, aero3d_server_*) against the plain
path and the oracle.

A served request is read back every check_every iterations from its
admission, through the same device reductions as the plain path's readback
(aero3d_readback), and its convergence rule is decided on the device. So each
result (iterations, residual, force coefficients, final state) must equal a
plain CfdHIP solve read back at the same iterations under the same rule, bit
for bit, whatever else is in the mix. The 30k-cell sphere's served state must
also match the oracle (the restatement of the reference CfdCPU) within 1e-5
per cell.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPECS = ["sphere:30000", "sphere:120000", "sphere:500000", "transport:30000", "transport:120000",
         "transport:500000", "interceptor:30000", "interceptor:120000", "interceptor:500000"]


def say(*a):
    print("[gpu-serve]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def built():
    import geometry
    say("building", len(SPECS), "geometries on the host")
    return geometry.build_cases(SPECS)


def freestream(M, alpha):
    a = alpha * math.pi / 180
    return [1, M * math.cos(a), 0, M * math.sin(a), 1 / (1.4 * 0.4) + 0.5 * M * M]


def setup(meta, M, alpha):
    F = freestream(M, alpha)
    fix = meta["fix"] if meta.get("fix") else [F, F, F]
    r = meta["ref"]
    return F, fix, (alpha, M, r["S"], r["xcg"], 0.0, r["c"])


def requests():
    """(geometry index, M, alpha, rule, res_drop, cap): both rules, caps not on the check grid,
    one residual-drop loose enough to converge early."""
    import aero3d
    out = []
    for k in range(len(SPECS)):
        M, alpha = 0.35 + 0.05 * k, -2 + 1.1 * k
        if k % 3 == 0:
            out.append((k, M, alpha, aero3d.RULE_REFERENCE, 0.0, 1500))
        elif k % 3 == 1:
            out.append((k, M, alpha, aero3d.RULE_RES_DROP, 1e-4, 430))      # capped between two readbacks
        else:
            out.append((k, M, alpha, aero3d.RULE_RES_DROP, 3e-2, 1000))     # converges on the residual drop
    return out


def plain_solve(g, geo, fl, meta, M, alpha, rule, res_drop, cap, every):
    import aero3d
    F, fix, ref = setup(meta, M, alpha)
    s = aero3d.CfdHIP(g, geo)
    s.opt["cfl"] = 0.5
    s.set_free(F); s.set_fix(fix); s.reset()
    s.set_force_list(fl)
    R = aero3d.ConvRule(rule, res_drop, cap)
    hist = s.solve(ref, R, every)
    Q = s.read()
    s.dispose()
    return hist, R, Q


def test_served_results_equal_plain_solves(built):
    import aero3d
    reqs = requests()
    every = 50
    # arena: the geometries plus room for about three of the largest cases, so later requests wait for memory
    probe = aero3d.Batch(arena_bytes=8 << 30, max_slots=len(SPECS))
    need = [probe.case_bytes(probe.add_geometry(g, geo, fl)) for (g, geo, fl, meta) in built]
    geo_bytes = probe.footprint()[0]
    probe.dispose()
    arena = geo_bytes + 3 * max(need) + (1 << 20)
    srv = aero3d.Server(arena_bytes=arena, max_slots=6, check_every=every)
    gids = [srv.batch.add_geometry(g, geo, fl) for (g, geo, fl, meta) in built]
    outs = {}
    order = [8, 0, 4, 2, 6, 1, 5, 3, 7]                 # large, small and middle grids interleaved
    done, submitted, ticks = {}, 0, 0
    while len(done) < len(reqs):
        # staggered arrivals: two requests every third tick
        if submitted < len(order) and ticks % 3 == 0:
            for k in order[submitted:submitted + 2]:
                gi, M, alpha, rule, rd, cap = reqs[k]
                F, fix, ref = setup(built[gi][3], M, alpha)
                outs[k] = np.empty((5, built[gi][3]["N"]), np.float32)
                srv.submit(gids[gi], F, fix, 0.5, 2, ref, rd, cap, srv.clock(), k, rule=rule, state_out=outs[k])
            submitted += 2
        a, p = srv.tick()
        ticks += 1
        for d in srv.poll():
            done[d["tag"]] = d
        assert ticks < 20000
    busy, gpu, cu = srv.stats()
    srv.dispose()
    say("served", len(done), "requests in", ticks, "ticks; device", round(gpu, 3), "s; cell-updates", cu)
    admits = sorted(d["t_admit"] for d in done.values())
    finishes = sorted(d["t_done"] for d in done.values())
    assert admits[-1] >= finishes[0], "every request was admitted at once: the arena did not force a wait"
    for k, (gi, M, alpha, rule, rd, cap) in enumerate(reqs):
        g, geo, fl, meta = built[gi]
        hist, R, Q = plain_solve(g, geo, fl, meta, M, alpha, rule, rd, cap, every)
        d, last = done[k], hist[-1]
        say(SPECS[gi], "rule", rule, "served", d["iters"], d["converged"], d["CL"], "plain", last["it"], R.conv, last["CL"])
        assert not d["failed"] and d["iters"] == last["it"] and bool(d["converged"]) == R.conv and d["nreads"] == len(hist)
        for key in ("res", "CL", "CD", "Cm"):
            assert d[key] == last[key], (SPECS[gi], key, d[key], last[key])
        assert d["res0"] == R.res0 or (math.isnan(d["res0"]) and math.isnan(R.res0))
        assert np.array_equal(outs[k].view(np.uint32), Q.view(np.uint32)), f"{SPECS[gi]}: served state differs from the plain solve"
        if rule == aero3d.RULE_RES_DROP and rd == 1e-4:
            assert d["iters"] == cap                    # the cap is reached exactly, between two readbacks
    assert any(done[k]["converged"] for k in range(len(reqs)))

    # the 30k sphere against the oracle at the served iterations
    import oracle3d as O
    from parity import rel_diff
    gi, M, alpha = reqs[0][:3]
    g, geo, fl, meta = built[gi]
    F, fix, ref = setup(meta, M, alpha)
    o = O.Oracle3D(g, geo, threads=16)
    o.set_free(F); o.set_fix(fix); o.set_opt(0.5, 2); o.reset(); o.step(done[0]["iters"])
    dd = rel_diff(outs[0], o.read(), geo["t"])
    say("sphere 30k served vs oracle after", done[0]["iters"], "iterations", dd)
    assert max(dd.values()) < 1e-5, dd


def test_server_rejects_bad_requests_and_never_stalls(built):
    """Bad requests are refused at submit; a request whose geometry is removed
    before admission comes back failed and the queue behind it is served."""
    import aero3d
    g, geo, fl, meta = built[0]
    F, fix, ref = setup(meta, 0.5, 1.0)
    srv = aero3d.Server(arena_bytes=1 << 30, max_slots=4, check_every=10)
    gid = srv.batch.add_geometry(g, geo, fl)
    for bad in ({"cfl": 0.0}, {"cfl": 12.0}, {"order": 3}, {"rule": 7}, {"max_iters": 0}, {"res_drop": 0.0}):
        args = dict(cfl=0.5, order=2, rule=aero3d.RULE_RES_DROP, max_iters=20, res_drop=1e-4)
        args.update(bad)
        with pytest.raises(aero3d.Aero3DError):
            srv.submit(gid, F, fix, args["cfl"], args["order"], ref, args["res_drop"], args["max_iters"], srv.clock(), 9,
                       rule=args["rule"])
    # one case resident, two queued behind it: the first queued one loses its geometry
    gid2 = srv.batch.add_geometry(g, geo, fl)
    srv.submit(gid, F, fix, 0.5, 2, ref, 1e-9, 40, srv.clock(), 1)
    srv.batch.set_zchunk(3)
    srv.tick()
    srv.submit(gid2, F, fix, 0.5, 2, ref, 1e-9, 20, srv.clock(), 2)
    srv.submit(gid, F, fix, 0.5, 2, ref, 1e-9, 20, srv.clock(), 3)
    srv.batch.remove_geometry(gid2)               # no case uses it yet
    done, ticks = {}, 0
    while len(done) < 3:
        srv.tick()
        ticks += 1
        for d in srv.poll():
            done[d["tag"]] = d
        assert ticks < 1000, "the queue stalled"
    srv.dispose()
    say("results", done)
    assert done[2]["failed"] == 1 and done[2]["iters"] == 0 and math.isnan(done[2]["CL"])
    assert done[1]["iters"] == 40 and done[3]["iters"] == 20 and not done[1]["failed"] and not done[3]["failed"]
