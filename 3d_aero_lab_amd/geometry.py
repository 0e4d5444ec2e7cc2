"""Python side of the JS host layer: runs `node js/cli.js geom` and loads its output.

The grid and cut-cell masks are built by the JS host (3d_aero_lab_amd/js),
which keeps the reference's double-precision classification order.
"""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(_HERE, "js", "cli.js")


def node_available():
    return shutil.which("node") is not None


def build_case(spec, outdir=None, timeout=900):
    """Build case `spec` (e.g. 'sphere', 'transport:1000000') -> (grid, geo, force_list, meta)."""
    own = outdir is None
    d = tempfile.mkdtemp(prefix="aero3d_geom_") if own else outdir
    try:
        subprocess.run(["node", CLI, "geom", spec, d], check=True, capture_output=True, timeout=timeout)
        return load_dir(d)
    finally:
        if own:
            shutil.rmtree(d, ignore_errors=True)


def load_dir(d):
    meta = json.load(open(os.path.join(d, "meta.json")))
    rd = lambda n, dt: np.fromfile(os.path.join(d, n + ".bin"), dtype=dt)
    grid = {"NX": meta["NX"], "NY": meta["NY"], "NZ": meta["NZ"], "fx": rd("fx", np.float64), "fy": rd("fy", np.float64),
            "fz": rd("fz", np.float64)}
    for ax in ("x", "y", "z"):
        f = grid["f" + ax]
        grid["d" + ax] = (f[1:] - f[:-1]).astype(np.float32)
        grid[ax + "c"] = (0.5 * (f[:-1] + f[1:])).astype(np.float32)
    geo = {"t": rd("t", np.uint8), "phi": rd("phi", np.float32), "ap": [rd(f"ap{k}", np.float32) for k in range(4)]}
    fl = {"q": rd("fl_q", np.int32), "cx": rd("fl_cx", np.float64), "cz": rd("fl_cz", np.float64),
          "x": rd("fl_x", np.float64), "z": rd("fl_z", np.float64)}
    return grid, geo, fl, meta


def build_cases(specs, timeout=1800):
    """Build several cases with one node process -> list of (grid, geo, force_list, meta)."""
    d = tempfile.mkdtemp(prefix="aero3d_geoms_")
    try:
        subprocess.run(["node", CLI, "geom-many", d, *specs], check=True, capture_output=True, timeout=timeout)
        return [load_dir(os.path.join(d, str(k))) for k in range(len(specs))]
    finally:
        shutil.rmtree(d, ignore_errors=True)


def mulberry32(seed):
    """Seeded uniform generator (same sequence as bodies.js mulberry32)."""
    a = seed & 0xFFFFFFFF

    def rnd():
        nonlocal a
        a = (a + 0x6D2B79F5) & 0xFFFFFFFF
        t = a
        t = ((t ^ (t >> 15)) * (t | 1)) & 0xFFFFFFFF
        t ^= (t + (((t ^ (t >> 7)) * (t | 61)) & 0xFFFFFFFF)) & 0xFFFFFFFF
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0
    return rnd
