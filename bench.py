#!/usr/bin/env python3
"""Benchmark of the plain 3D Euler cut-cell path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: the 40 m transport (airliner layout, AR 9.5 wing, 25 deg sweep,
taper 0.3, 12 % sections, tail at 15 % of the wing area, two nacelle
fixed-state groups) on the 1.0M-cell stretched half-domain grid, M 0.78,
alpha 2 deg, CFL 0.5, FP32 state. A step is one SSP-RK2 iteration (two stage
kernels) over all cells; state, masks and grid are resident in HBM before the
timed region. Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d_aero_lab_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL_UPDATE = 142      # 61 B stage 1 + 81 B stage 2 (FP32 SoA state, t u8, phi f32, ap[4] f32)
WORKLOAD = "transport:1000000"


def cpu_baseline(grid, geo, F, fix, budget_s=15.0):
    """The oracle (C restatement of the reference CfdCPU) on the host cores, on
    a bounded number of iterations of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle3d
    threads = min(os.cpu_count() or 1, 16)
    o = oracle3d.Oracle3D(grid, geo, threads=threads)
    o.set_free(F); o.set_fix(fix); o.set_opt(0.5, 2); o.reset()
    t0 = time.perf_counter(); o.step(1); t1 = time.perf_counter()
    n = max(1, min(200, int(budget_s / max(t1 - t0, 1e-3))))
    t0 = time.perf_counter(); o.step(n); dt = time.perf_counter() - t0
    N = len(geo["t"])
    return {"value": N * n / dt, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} SSP-RK2 iterations of the same 1.0M-cell transport (after 1 untimed), OpenMP over cells"}


def copy_probe(nbytes=1 << 30):
    """A plain device-to-device copy of `nbytes` (reads nbytes, writes nbytes)."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": 3}), flush=True)


def freestream(M, alpha):
    import math
    a = alpha * math.pi / 180
    return [1, M * math.cos(a), 0, M * math.sin(a), 1 / (1.4 * 0.4) + 0.5 * M * M]


def case_setup(meta, M, alpha):
    """Freestream, fixed states and force-coefficient reference of one built case."""
    F = freestream(M, alpha)
    fix = meta["fix"] if meta.get("fix") else [F, F, F]
    r = meta["ref"]
    return F, fix, (alpha, M, r["S"], r["xcg"], 0.0, r["c"])


def batched_workload(seed=20261203, n=64):
    """64 cases: sphere or transport (p = 1/2), cells ~ U[30k, 1.0M], M ~ U[0.3, 0.85],
    alpha ~ U[-2, 8] deg, tail deflection ~ U[-5, 5] deg, all from one mulberry32 seed."""
    import geometry
    rnd = geometry.mulberry32(seed)
    out = []
    for _ in range(n):
        body = "sphere" if rnd() < 0.5 else "transport"
        cells = int(30000 + rnd() * 970000)
        M, alpha, delta = 0.3 + 0.55 * rnd(), -2 + 10 * rnd(), -5 + 10 * rnd()
        out.append((f"sphere:{cells}" if body == "sphere" else f"transport:{cells}:{delta:.6f}", M, alpha))
    return out


def run_converged_plain(built, conds, res_drop, max_iters, check_every, tile=None):
    """Each case alone on the plain path (tile: stage-kernel shape, None: the solver's default for the grid, the
    width rule the batched form applies per case), one after another, read back every check_every iterations
    (aero3d_readback) under the residual-drop rule (res/res0 <= res_drop, or max_iters), the server's
    rule. Returns the wall time, the iterations and the final (CL, CD, Cm) of each case."""
    import aero3d
    t0 = time.perf_counter()
    iters, coefs = [], []
    for (grid, geo, fl, meta), (M, alpha) in zip(built, conds):
        F, fix, fref = case_setup(meta, M, alpha)
        s = aero3d.CfdHIP(grid, geo)
        s.opt["cfl"] = 0.5
        if tile is not None:
            s.set_tile(*tile)
        s.set_free(F); s.set_fix(fix); s.reset()
        s.set_force_list(fl)
        hist = s.solve(fref, aero3d.ConvRule(aero3d.RULE_RES_DROP, res_drop, max_iters), check_every)
        iters.append(hist[-1]["it"])
        coefs.append((hist[-1]["CL"], hist[-1]["CD"], hist[-1]["Cm"]))
        s.dispose()
    return time.perf_counter() - t0, iters, coefs


def coef_rel_diff(done, coefs):
    """Largest relative CL / CD difference between server results (by tag) and plain coefficients."""
    worst = 0.0
    for d in done:
        cl, cd, _ = coefs[d["tag"]]
        for a, b in ((d["CL"], cl), (d["CD"], cd)):
            worst = max(worst, abs(a - b) / max(abs(b), 1e-12))
    return worst


def run_server(built, reqs, res_drop, max_iters, check_every, zc, arena_bytes, max_slots=256, group_cells=1 << 40, persistent=0):
    """Submit requests (geometry index, M, alpha, arrival s) to the continuous-batching server as their
    arrival times come; tick until all are done. Returns results and the wall time."""
    import aero3d
    srv = aero3d.Server(arena_bytes=arena_bytes, max_slots=max_slots, check_every=check_every)
    srv.batch.set_zchunk(zc)
    srv.batch.set_persistent(persistent)
    srv.batch.set_group_cells(group_cells)
    gids = [srv.batch.add_geometry(g, geo, fl) for (g, geo, fl, meta) in built]
    t_start = srv.clock()
    pending = sorted(range(len(reqs)), key=lambda k: reqs[k][3])
    done, nxt = [], 0
    while len(done) < len(reqs):
        now = srv.clock() - t_start
        while nxt < len(pending) and reqs[pending[nxt]][3] <= now:
            k = pending[nxt]
            gi, M, alpha, arr = reqs[k]
            F, fix, fref = case_setup(built[gi][3], M, alpha)
            srv.submit(gids[gi], F, fix, 0.5, 2, fref, res_drop, max_iters, t_start + arr, k)
            nxt += 1
        a, p = srv.tick()
        done += srv.poll()
        if a == 0 and p == 0 and nxt < len(pending):
            time.sleep(max(0.0, reqs[pending[nxt]][3] - (srv.clock() - t_start)))
    wall = srv.clock() - t_start
    busy, gpu, cu = srv.stats()
    srv.dispose()
    return done, wall, gpu, cu


def batched_leg(res_drop=1e-4, max_iters=2000, check_every=50, config=None):
    """Converged solves/s of the 64-case batch (one launch per stage, graph-replayed) against the same cases
    solved one at a time on the plain path. config: (zchunk, group_cells, persistent), or None to tune them."""
    import aero3d
    import geometry
    wl = batched_workload()
    print(f"[bench] batched: building {len(wl)} geometries", file=sys.stderr, flush=True)
    built = geometry.build_cases([w[0] for w in wl])
    conds = [(w[1], w[2]) for w in wl]
    cells = sum(b[3]["N"] for b in built)
    if config is None:
        # autotune the z chunk and the case-group size on a throwaway batch holding every case
        tb = aero3d.Batch(arena_bytes=int(cells * 110 + (2 << 30)), max_slots=128)
        for (g, geo, fl, meta), (M, alpha) in zip(built, conds):
            F, fix, _ = case_setup(meta, M, alpha)
            tb.add(tb.add_geometry(g, geo), F, fix, 0.5)
        best = None
        for pers in (0, 1):
            tb.set_persistent(pers)
            for zc in (0, 3, 4, 6, 8):          # 0: from the load
                for gc in (1_000_000, 2_000_000, 4_000_000, 1 << 40):
                    tb.set_zchunk(zc)
                    tb.set_group_cells(gc)
                    tb.time_steps(3)
                    ms = tb.time_steps(3)
                    if best is None or ms < best[1]:
                        best = (zc, ms, gc, pers)
        tb.dispose()
        zc, group, pers = best[0], best[2], best[3]
    else:
        zc, group, pers = config
    print(f"[bench] batched: {cells} cells, zchunk {zc}, group {group}, persistent {pers}; server run", file=sys.stderr, flush=True)
    reqs = [(k, M, alpha, 0.0) for k, (M, alpha) in enumerate(conds)]
    done, wall_b, gpu_b, cu = run_server(built, reqs, res_drop, max_iters, check_every, zc, arena_bytes=int(cells * 110 + (2 << 30)),
                                         group_cells=group, persistent=pers)
    print("[bench] batched: plain path, one case at a time", file=sys.stderr, flush=True)
    wall_p, iters_p, coefs_p = run_converged_plain(built, conds, res_drop, max_iters, check_every)
    it_b = {d["tag"]: d["iters"] for d in done}
    same = all(it_b[k] == iters_p[k] for k in range(len(conds)))
    dcoef = coef_rel_diff(done, coefs_p)
    return {"cases": len(conds), "cells": cells, "solves_per_s": len(conds) / wall_b, "wall_s": wall_b,
            "cell_updates_per_s": cu / wall_b, "plain_solves_per_s": len(conds) / wall_p, "plain_wall_s": wall_p,
            "speedup_vs_plain": wall_p / wall_b, "converged": sum(d["converged"] for d in done),
            "iters_mean": sum(it_b.values()) / len(it_b), "same_iterations_as_plain": same, "cl_cd_max_rel_diff_vs_plain": dcoef,
            "gpu_busy_frac": gpu_b / wall_b,
            "zchunk": zc, "group_cells": group, "launch": "persistent" if pers else "one block per work item",
            "criterion": f"res/res0 <= {res_drop} (density residual every {check_every} iterations) or {max_iters} iterations"}


def serve_workload(seed=20261204, n=10000):
    """Poisson arrivals (mean 20 req/s); body sphere / transport / swept-delta single fuselage (1/3 each);
    level 30k / 120k / 500k cells (0.5 / 0.3 / 0.2); M ~ U[0.3, 0.85], alpha ~ U[-2, 8] deg."""
    import geometry
    rnd = geometry.mulberry32(seed)
    t, out = 0.0, []
    for _ in range(n):
        t += -math.log(1 - rnd()) * 0.05
        b = int(rnd() * 3)
        u = rnd()
        lv = 0 if u < 0.5 else (1 if u < 0.8 else 2)
        M, alpha = 0.3 + 0.55 * rnd(), -2 + 10 * rnd()
        out.append((3 * b + lv, M, alpha, t))
    return out


SERVE_SPECS = ["sphere:30000", "sphere:120000", "sphere:500000", "transport:30000", "transport:120000",
               "transport:500000", "interceptor:30000", "interceptor:120000", "interceptor:500000"]


def serve_leg(n_requests=200, res_drop=1e-4, max_iters=2000, check_every=50):
    """The served form on a sample of the seeded request stream (the full 10k-request stream is a long run of
    its own: js/serve_stream.js), its capacity with every request arriving at once, and the capacity per grid
    level against the plain path (the gain of batching by grid size)."""
    import geometry
    reqs = serve_workload()[:n_requests]
    print(f"[bench] serve: {len(reqs)} requests over {reqs[-1][3]:.1f} s of arrivals", file=sys.stderr, flush=True)
    built = geometry.build_cases(SERVE_SPECS)
    done, wall, gpu, cu = run_server(built, reqs, res_drop, max_iters, check_every, 0, arena_bytes=64 << 30)
    lat = sorted(d["t_done"] - d["arrival"] for d in done)
    pct = lambda p: lat[min(len(lat) - 1, int(round(p / 100 * (len(lat) - 1))))]
    # the same requests one at a time on the plain path (arrival times ignored: its best case)
    print("[bench] serve: plain path, one request at a time", file=sys.stderr, flush=True)
    wall_p, iters_p, coefs_p = run_converged_plain([built[r[0]] for r in reqs], [(r[1], r[2]) for r in reqs], res_drop, max_iters,
                                                   check_every)
    it_s = {d["tag"]: d["iters"] for d in done}
    # capacity: the same requests all arriving at once (the stream above is arrival-bound at ~20 req/s)
    print("[bench] serve: capacity run, all requests at t = 0", file=sys.stderr, flush=True)
    done_c, wall_c, gpu_c, cu_c = run_server(built, [(r[0], r[1], r[2], 0.0) for r in reqs], res_drop, max_iters, check_every, 0,
                                             arena_bytes=64 << 30)
    # capacity per grid level: the three bodies at one level, several requests each, all at once, against the plain path
    print("[bench] serve: capacity per grid level", file=sys.stderr, flush=True)
    levels = {}
    rnd = geometry.mulberry32(20261206)
    for cells, per_body in ((30000, 8), (120000, 4), (500000, 2), (1000000, 1)):
        specs = [f"sphere:{cells}", f"transport:{cells}", f"interceptor:{cells}"]
        lvl_built = [built[SERVE_SPECS.index(sp)] for sp in specs] if all(sp in SERVE_SPECS for sp in specs) else geometry.build_cases(specs)
        lreq = [(b, 0.3 + 0.55 * rnd(), -2 + 10 * rnd(), 0.0) for b in range(3) for _ in range(per_body)]
        dl, wl, gl, cul = run_server(lvl_built, lreq, res_drop, max_iters, check_every, 0, arena_bytes=64 << 30)
        wp, ip, _ = run_converged_plain([lvl_built[r[0]] for r in lreq], [(r[1], r[2]) for r in lreq], res_drop, max_iters,
                                        check_every)
        it_l = {d["tag"]: d["iters"] for d in dl}
        levels[str(cells)] = {"requests": len(lreq), "served_solves_per_s": len(lreq) / wl, "plain_solves_per_s": len(lreq) / wp,
                              "speedup_vs_plain": wp / wl, "gpu_busy_frac": gl / wl, "served_cell_updates_per_s": cul / wl,
                              "same_iterations_as_plain": all(it_l[k] == ip[k] for k in range(len(lreq)))}
        print(f"[bench] serve: level {cells}: {levels[str(cells)]}", file=sys.stderr, flush=True)
    return {"requests": len(done), "sample": f"first {n_requests} of the 10k-request stream (seed 20261204); "
                                             "the whole stream: js/serve_stream.js",
            "capacity_solves_per_s": len(done_c) / wall_c, "capacity_cell_updates_per_s": cu_c / wall_c,
            "capacity_speedup_vs_plain": wall_p / wall_c, "capacity_gpu_busy_frac": gpu_c / wall_c,
            "solves_per_s": len(done) / wall, "wall_s": wall, "p50_s": pct(50), "p95_s": pct(95), "p99_s": pct(99),
            "gpu_busy_frac": gpu / wall, "cell_updates_per_s": cu / wall, "converged": sum(d["converged"] for d in done),
            "plain_solves_per_s": len(reqs) / wall_p, "plain_wall_s": wall_p, "speedup_vs_plain": wall_p / wall,
            "same_iterations_as_plain": all(it_s[k] == iters_p[k] for k in range(len(reqs))),
            "cl_cd_max_rel_diff_vs_plain": coef_rel_diff(done, coefs_p),
            "capacity_by_grid_level": levels,
            "criterion": f"res/res0 <= {res_drop} or {max_iters} iterations, readback every {check_every} (the first takes the snapshot)"}


AXI_CASE = os.path.join(ROOT, "tests", "golden", "axi_tj_medium")


def axi_leg(iters=14000):
    """Turbojet axisymmetric run (C5): the reference setup's medium grid (120,600 cells), 14,000 SSP-RK2
    iterations with the Tt4 control every 250 iterations, thrust and station averages (js/axi_run.js)."""
    import subprocess
    print(f"[bench] axi: TJ medium, {iters} iterations", file=sys.stderr, flush=True)
    r = subprocess.run(["node", os.path.join(ROOT, "3d_aero_lab_amd", "js", "axi_run.js"), AXI_CASE, str(iters), "2000"],
                       capture_output=True, text=True, timeout=600, check=True)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out["cell_updates_per_s"] = out["cells"] * iters / out["step_s"]
    out["roofline_frac"] = out["cell_updates_per_s"] * 182 / (HBM_PEAK_GBS * 1e9)
    out["thrust_kN"] = out["thrust"]["F"] / 1000
    out["published_thrust_kN"] = 56.9
    # the application's convergence loop with and without grid sequencing (js/axi_seq.js;
    # readbacks on the device monitor, the state carried between grids on the device)
    seq = {}
    coarse_dir = os.path.join(ROOT, "tests", "golden", "axi_tj_coarse")
    for mode, coarse in (("sequenced", coarse_dir), ("direct", "-")):
        r = subprocess.run(["node", os.path.join(ROOT, "3d_aero_lab_amd", "js", "axi_seq.js"), coarse, AXI_CASE, "250", "60000"],
                           capture_output=True, text=True, timeout=600, check=True)
        d = json.loads(r.stdout.strip().splitlines()[-1])
        seq[mode] = {"grids": [{k: g[k] for k in ("W", "H", "iters")} for g in d["grids"]], "thrust_kN": d["F"] / 1000,
                     "residual": d["res"], "wall_s": d["wall_s"]}
    out["convergence_loop"] = seq
    return out


def axi_cpu_baseline(budget_s=10.0):
    """The axisymmetric oracle (C restatement of the reference CPUSolver, 1 thread) on the same medium case."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_axi
    from axi_case import load
    g, z, Q0, comps, xf, rf = load("axi_tj_medium")
    s = g.s
    o = oracle_axi.OracleAxi(s["W"], s["H"], xf, rf, z, comps, s["fs"], list(Q0), dict(s["opt"], hf=s["hf"]))
    t0 = time.perf_counter(); o.step(1); t1 = time.perf_counter()
    n = max(1, min(500, int(budget_s / max(t1 - t0, 1e-3))))
    t0 = time.perf_counter(); o.step(n); dt = time.perf_counter() - t0
    return {"value": s["N"] * n / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{n} iterations of the TJ medium case (after 1 untimed), single thread"}


def sphere_leg():
    """C1: the 48 x 32 x 24 sphere case (radius 4.8 cells, M 0.4, CFL 0.5) read back every 50 iterations until
    the reference controller's convergence rule (src/air_cfd.js:573-576) ends it: iterations and drag
    coefficient C_d (S = pi R^2), beside the reference's own run of the same case (tests/golden/readback_sphere)
    and the README's figure for radius 4.8 cells (README.md:99, grid not stated)."""
    import geometry
    import aero3d
    grid, geo, fl, meta = geometry.build_case("sphere")
    R = 4.8
    M = 0.4
    F = [1, M, 0, 0, 1 / (1.4 * 0.4) + 0.5 * M * M]
    s = aero3d.CfdHIP(grid, geo)
    s.opt["cfl"] = 0.5
    s.set_free(F); s.set_fix([F, F, F]); s.reset(); s.set_force_list(fl)
    t0 = time.perf_counter()
    hist = s.solve((0.0, M, math.pi * R * R, 0.0, 0.0, 1.0), aero3d.ConvRule(aero3d.RULE_REFERENCE, cap=30000), 50, 2 * R * 1.12)
    wall = time.perf_counter() - t0
    s.dispose()
    last = hist[-1]
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "readback_sphere", "manifest.json")))["scalars"]["reads"][-1]
    return {"cells": meta["N"], "converged_at": last["it"], "Cd": last["CD"], "CDi": last["CDi"], "Mmax": last["Mmax"], "wall_s": wall,
            "reference_run_same_grid": {"converged_at": ref["it"], "Cd": ref["CD"]}, "readme_Cd_radius_4_8": 0.56,
            "rule": "reference controller: > 6 readbacks, res/res0 < 2e-3, CL spread of the last 4 < 2e-3; readback every 50"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--legs", default="plain,sphere,batched,serve,axi", help="comma list of plain, sphere, batched, serve, axi")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--batch-config", default="auto",
                    help="batched leg: 'zchunk,group_cells,persistent' (0 = from the load, group cells, 0 / 1) or 'auto' (tuner)")
    ap.add_argument("--tile", default="auto", help="stage-kernel shape 'ty,zchunk' or 'auto' (autotuner)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of known size (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: one GPU only; a solve would shard by z-slabs with a 2-plane halo exchange per stage, "
              "but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        return copy_probe()
    import geometry
    import aero3d
    grid, geo, fl, meta = geometry.build_case(WORKLOAD)
    import math
    M, alpha = 0.78, 2.0
    al = alpha * math.pi / 180
    F = [1, M * math.cos(al), 0, M * math.sin(al), 1 / (1.4 * 0.4) + 0.5 * M * M]
    fix = meta["fix"]
    s = aero3d.CfdHIP(grid, geo)
    s.opt["cfl"] = 0.5
    s.set_free(F); s.set_fix(fix); s.reset()
    s.set_force_list(fl)
    print(f"[bench] grid {meta['NX']}x{meta['NY']}x{meta['NZ']} = {meta['N']} cells, warmup {a.warmup}", file=sys.stderr, flush=True)
    if a.tile == "auto":
        ty, zc, _ = s.autotune(10)
    else:
        ty, zc = (int(x) for x in a.tile.split(","))
        s.set_tile(ty, zc)
    s.step(a.warmup)
    s.sync()
    t0 = time.perf_counter()
    ms = s.time_steps(a.steps)          # HIP events on the solver's stream around the K iterations
    wall = time.perf_counter() - t0
    N = meta["N"]
    sec = ms / 1e3
    value = N * a.steps / sec
    launch_s = sec / (2 * a.steps)
    achieved = BYTES_PER_CELL_UPDATE / 2 * N / launch_s / 1e9
    r = meta["ref"]
    f = s.forces(alpha, M, r["S"], r["xcg"], 0, r["c"])
    # C2: the same solve continued to 2000 iterations; its coefficients, Trefftz CDi and Mach maximum
    it_timed = s.it
    if s.it < 2000:
        s.step(2000 - s.it)
    rb = s.readback(alpha, M, r["S"], r["xcg"], 0.0, r["c"], meta["xT"] if meta.get("xT") is not None else float("nan"))
    c2 = {"it": rb["it"], "CL": rb["CL"], "CD": rb["CD"], "Cm": rb["Cm"], "CDi": rb["CDi"], "Mmax": rb["Mmax"]}
    traffic = None
    try:     # HBM bytes per launch measured by rocprofv3 counters for this kernel shape (profiles/pmc_traffic.json)
        t = json.load(open(os.path.join(ROOT, "profiles", "pmc_traffic.json")))["plain_transport_1m"]
        if t["tile"] == [64, ty, zc]:
            traffic = t["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        pass
    out = {
        "metric": "cell-updates/s", "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "plain 3D Euler, transport body, 1.0M-cell stretched half domain, M0.78, alpha 2, CFL 0.5",
                   "cells": N, "grid": [meta["NX"], meta["NY"], meta["NZ"]], "parallelism": "single GPU",
                   "tile": [64, ty, zc]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "wall_s": wall, "it_timed": it_timed, "CL_at_it_timed": f["CL"], "CD_at_it_timed": f["CD"],
        "transport_2000_iterations": c2,
    }
    s.dispose()
    legs = a.legs.split(",")
    if "sphere" in legs:
        out["sphere"] = sphere_leg()
    if "batched" in legs:
        out["batched"] = batched_leg(config=None if a.batch_config == "auto" else tuple(int(x) for x in a.batch_config.split(",")))
    if "serve" in legs:
        out["served"] = serve_leg(a.serve_requests)
    if "axi" in legs:
        out["axisymmetric"] = axi_leg()
        if not a.no_cpu:
            out["axisymmetric"]["cpu_baseline"] = axi_cpu_baseline()
    if not a.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(grid, geo, F, fix)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
