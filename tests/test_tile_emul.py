"""CPU build of the tiled stage kernel (same device header, g++, one OS thread
per GPU thread, AddressSanitizer on every global array) against the oracle.
Checks the tile / halo / z-chunk indexing before any GPU run."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EMUL = os.path.join(HERE, "cpu_emul")


@pytest.fixture(scope="module")
def emul_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("emul") / "libtile_emul.so")
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fPIC", "-shared", "-fsanitize=address", "-fno-omit-frame-pointer",
                    "-I", EMUL, "-o", out, os.path.join(EMUL, "tile_emul.cpp"), "-lpthread"], check=True)
    return out


@pytest.mark.parametrize("case,iters,ty,zc", [("sphere", 2, 4, 5), ("transport60k", 2, 8, 3), ("interceptor", 1, 8, 40),
                                               ("sphere", 2, 32, 4)])
def test_tile_kernel_cpu_build_matches_oracle(emul_lib, case, iters, ty, zc):
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(EMUL, "run_emul.py"), emul_lib, case, str(iters), str(ty), str(zc)],
                       capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    import json
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert max(d.values()) < 1e-5, d


# one OS thread per emulated GPU thread makes each run cost seconds: a subset of
# tests/synth_cases.py SIZES (single cells, sub-tile, just past a tile, tall and
# thin), each with a z chunk shorter and one longer than the grid; two at Mach 2
# along +x and -x (the HLLC's supersonic branches on most faces)
SYNTH = [("synth:1:2:1", 8, 3), ("synth:2:2:3", 4, 1), ("synth:3:5:2", 8, 1), ("synth:65:9:3", 8, 2),
         ("synth:7:40:2", 4, 3), ("synth:200:2:1", 8, 3), ("synth:33:9:3", 32, 2), ("synth:1:2:1", 32, 3),
         ("synth:33:9:3@2.0", 32, 2), ("synth:65:9:3@-2.0", 8, 2)]


@pytest.mark.parametrize("case,ty,zc", SYNTH)
def test_tile_kernel_cpu_build_small_grids(emul_lib, case, ty, zc):
    """Grids of one to a few tiles, narrower and shorter than a tile and a z
    chunk (tests/synth_cases.py), under AddressSanitizer: every access in
    bounds and the oracle reproduced after 3 iterations."""
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(EMUL, "run_emul.py"), emul_lib, case, "3", str(ty), str(zc)],
                       capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    import json
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert max(d.values()) < 1e-5, d


def test_axi_kernels_cpu_build_matches_oracle(tmp_path):
    """The axisymmetric face + cell kernels and the fused tile kernel under
    AddressSanitizer: the oracle reproduced after 2 iterations, and the fused
    kernel (LDS tiles, one OS thread per GPU thread) bit-identical to the
    face + cell kernels (fused_same = 0)."""
    out = str(tmp_path / "libaxi_emul.so")
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fPIC", "-shared", "-fsanitize=address", "-fno-omit-frame-pointer",
                    "-I", EMUL, "-o", out, os.path.join(EMUL, "axi_emul.cpp")], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(EMUL, "run_axi_emul.py"), out, "axi_tj_xcoarse", "2", "fused"],
                       capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stderr[-3000:]
    import json
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert max(d.values()) < 1e-5, d
