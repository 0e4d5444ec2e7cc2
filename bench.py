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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tile", default="auto", help="stage-kernel shape 'ty,zchunk' or 'auto' (autotuner)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain device copy of known size (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only; the solve would shard by z-slabs with a 2-cell halo "
              "(see DESIGN.md) but no multi-GPU path is built", file=sys.stderr)
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
        "wall_s": wall, "it": s.it, "CL": f["CL"], "CD": f["CD"],
    }
    if not a.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(grid, geo, F, fix)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
