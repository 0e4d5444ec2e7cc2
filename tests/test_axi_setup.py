"""Axisymmetric setup restatement (3d_aero_lab_amd/js/cfd_grid.js: buildCFDGrid,
buildZones, updateHeatZones, streamMassFlows, compParams; src/cfd_grid.js:11-276)
against the reference's own setups: grid faces, every zone array and the
component constants byte-identical (sha256) for the turbojet, the high- and
low-bypass turbofans at all four grid levels (tests/golden/axi_setups) and for
the five setups the solver fixtures carry."""
import json
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = ["axi_setups", "axi_tj_xcoarse", "axi_tj_medium", "axi_hbtf_xcoarse", "axi_hbtf_medium", "axi_lbtf_xcoarse"]


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
@pytest.mark.parametrize("case", CASES)
def test_axi_setup_byte_identical(case):
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "axi_setup_check.js"), os.path.join(ROOT, "tests", "golden", case)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout)
    assert res
    bad = {name: [k for k, ok in v.items() if not ok] for name, v in res.items() if not all(v.values())}
    assert not bad, bad


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_checkpoint_format_matches_reference(tmp_path):
    """Long-run checkpoint files (save / load, tools/cfdlib.js:16-21): the TJ
    xcoarse state after 10 iterations written by js/cfd_axi.js has the same
    bytes as the reference's own save() of it, its constants the same JSON
    text; loading returns the planes exactly and ignores a file sized for
    another grid, as the reference's loader does."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "axi_ckpt_check.js"), os.path.join(ROOT, "tests", "golden"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout)
    assert all(res.values()), res


def test_pool_setups_byte_identical():
    """The worker-thread setup (GeomPool.axiSetup: buildZones' column pass in
    slabs of 1, 7 and the default column count over 3 workers + the caller)
    gives the reference's grids, zones and constants for all 12 setups."""
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "geom_pool_axi_check.js"), "3",
                        os.path.join(ROOT, "tests", "golden", "axi_setups")], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert len(res) == 36
    bad = {k: v for k, v in res.items() if v is not True}
    assert not bad, bad

