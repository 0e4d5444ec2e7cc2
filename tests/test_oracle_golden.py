"""The CPU oracle against the reference's own outputs (tests/golden).

The oracle restates CfdCPU (src/air_cfd.js:157-259) in C with the same double
arithmetic and Float32 storage, so it must reproduce the reference bit for bit.
"""
import numpy as np
import pytest

import oracle3d as O
from golden_io import Golden


def run_oracle(g, tags_iters, order=2):
    grid, geo = g.grid(), g.geo()
    o = O.Oracle3D(grid, geo)
    F = g.s["F"]
    o.set_free(F)
    o.set_fix(g.s.get("fix", [F, F, F]))
    o.set_opt(g.s["cfl"], order)
    o.reset()
    done = 0
    out = {}
    for tag, it in tags_iters:
        o.step(it - done)
        done = it
        out[tag] = o.read()
    return o, out


@pytest.mark.parametrize("case,order,tags", [
    ("sphere", 2, [("it1", 1), ("it10", 10), ("it200", 200)]),
    ("sphere_o1", 1, [("it10", 10)]),
    ("interceptor", 2, [("it20", 20)]),
    ("transport60k", 2, [("it2", 2), ("it100", 100)]),
])
def test_oracle_bit_exact(case, order, tags):
    g = Golden(case)
    if case == "sphere_o1":
        base = Golden("sphere")
        g.grid, g.geo = base.grid, base.geo
        g.s = dict(base.s, **{k: v for k, v in g.s.items() if k != "hist"})
    _, out = run_oracle(g, tags, order)
    for tag, _ in tags:
        ref = g.state(tag)
        assert np.array_equal(out[tag].view(np.uint32), ref.view(np.uint32)), f"{case} {tag}: oracle differs from reference"


def test_oracle_forces_history_sphere():
    g = Golden("sphere")
    grid, geo, fl = g.grid(), g.geo(), g.force_list()
    o = O.Oracle3D(grid, geo)
    F = g.s["F"]
    o.set_free(F); o.set_fix([F, F, F]); o.set_opt(0.5, 2); o.reset()
    done = 0
    for h in g.s["hist"]:
        o.step(h["it"] - done); done = h["it"]
        f = O.forces(o.read(), fl, 0, g.s["M"], g.s["S"], 0, 0, 1)
        assert f["CD"] == h["CD"] and f["CL"] == h["CL"] and f["Cm"] == h["Cm"]


def test_oracle_fighter_readme_row():
    """README.md:98: fighter 30k cells, M0.6, alpha 3, delta 3, 200 iterations: CL 0.1927, CD 0.04406."""
    g = Golden("fighter30k")
    grid, geo, fl = g.grid(), g.geo(), g.force_list()
    o = O.Oracle3D(grid, geo, threads=4)
    F = g.s["F"]
    o.set_free(F); o.set_fix([F, F, F]); o.set_opt(0.7, 2); o.reset()
    o.step(200)
    Q = o.read()
    assert np.array_equal(Q.view(np.uint32), g.state("it200").view(np.uint32))
    r = g.s["ref"]
    f = O.forces(Q, fl, g.s["alpha"], g.s["M"], r["S"], r["xcg"], 0, r["c"])
    assert f["CL"] == g.s["hist"][-1]["CL"] and f["CD"] == g.s["hist"][-1]["CD"]
    assert round(f["CL"], 4) == 0.1927 and round(f["CD"], 5) == 0.04406
    cdi = O.trefftz(grid, geo["t"], Q, g.s["xT"], g.s["alpha"], g.s["M"], r["S"])
    assert cdi == g.s["CDi"]


def test_oracle_threads_do_not_change_results():
    g = Golden("interceptor")
    F = g.s["F"]
    res = []
    for th in (1, 3):
        o = O.Oracle3D(g.grid(), g.geo(), threads=th)
        o.set_free(F); o.set_fix([F, F, F]); o.set_opt(0.5, 2); o.reset(); o.step(3)
        res.append(o.read())
    assert np.array_equal(res[0], res[1])
