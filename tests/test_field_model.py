"""The engine field model (FieldModel.sample, src/field.js:296-488), the
axisymmetric initial state (initialState with streamState1D,
src/cfd_grid.js:242-305) and the 3D fixed states (CFD3.applyCond,
src/air_cfd.js:521-538) of 3d_aero_lab_amd/csrc/field.hpp, checked on the CPU
before any device run: a g++ build of the same device functions under
AddressSanitizer against the reference's own outputs (tests/golden/axi_field
from make_golden_axi.js, tests/golden/applycond from make_golden.js), and the
table builders of Node and Python byte for byte.

Bars: regions (outside / solid / which stream) identical at all 15,000 seeded
points of five models (three engines, two with the afterburner lit); values
within 1e-13 relative (the device's pow / exp / sin / cos / tan / log against
V8's, measured <= 1.6e-15 on this CPU build); initial states within 1e-12 per
cell, uniform starts bit for bit; corrupted tables refused before any
evaluation."""
import hashlib
import json
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMUL = os.path.join(HERE, "cpu_emul")


def test_field_model_cpu_build_matches_reference(tmp_path):
    lib = str(tmp_path / "libfield_emul.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-fsanitize=address", "-fno-omit-frame-pointer",
                    "-ffp-contract=off", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "3d_aero_lab_amd", "csrc"),
                    "-o", lib, os.path.join(EMUL, "field_emul.cpp")], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(EMUL, "run_field_emul.py"), lib], capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    for name, s in d["samples"].items():
        print(name, s)
        assert s["region_mismatch"] == 0 and s["max_rel"] < 1e-13, (name, s)
    assert set(d["samples"]["lbtf"]["regions"]) == {0, 1, 2, 3, 4, 5}     # outside, solid and all four streams
    for case, s in d["initial"].items():
        print(case, s)
        assert s["max_rel"] < 1e-12 and s["uniform_exact"], (case, s)
    assert d["sound_table_accepted"]
    assert all(d["rejected"].values()), d["rejected"]


@pytest.mark.skipif(shutil.which("node") is None, reason="node is not installed")
def test_field_table_js_equals_python():
    import aero3d
    from golden_io import Golden
    r = subprocess.run(["node", os.path.join(HERE, "js", "field_check.js"), "tables"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    js = json.loads(r.stdout)
    py = {k: hashlib.sha256(aero3d.field_table(d).tobytes()).hexdigest() for k, d in Golden("axi_field").s["models"].items()}
    py.update({"applycond_" + k: hashlib.sha256(aero3d.field_table(c["model"]).tobytes()).hexdigest()
               for k, c in Golden("applycond").s["cases"].items()})
    assert js == py
