"""Invariants of the scheme on the oracle (src/air_cfd.js:198-251), on grids the
reference's fixtures do not cover (tests/synth_cases.py):

* freestream preservation: on a grid with no body every face sees the same
  state on both sides and F(U, U) = F(U), so the state stays at the freestream
  (to round-off: the ghost cells' primitives come from the double freestream,
  the cells' from the FP32 state);
* gas at rest stays at rest around a cut-cell body: the wall pressure source
  (src/air_cfd.js:225) cancels the pressure flux through the partly open
  faces, and a baffle carries no mass at u = 0 (src/air_cfd.js:226-233).

A sign error in the wall source gives |rho u| ~ 5 after 20 iterations on the
65 x 9 x 5 grid, against 6e-9 for the oracle."""
import numpy as np
import pytest

import oracle3d as O
from synth_cases import synth_case

GRIDS = [(65, 9, 5), (7, 40, 2), (3, 5, 2)]


def run(n, iters, **kw):
    grid, geo, F, fix = synth_case(*n, seed=1, **kw)
    o = O.Oracle3D(grid, geo)
    o.set_free(F); o.set_fix(fix); o.set_opt(0.5, 2); o.reset(); o.step(iters)
    return o.read().astype(np.float64), geo, F


@pytest.mark.parametrize("n", GRIDS, ids=lambda n: "x".join(map(str, n)))
def test_freestream_is_preserved(n):
    Q, geo, F = run(n, 200, body=False)
    F32 = np.array(F, np.float32).astype(np.float64)
    dev = max(np.abs(Q[c] - F32[c]).max() for c in range(5)) / max(abs(f) for f in F)
    assert dev < 1e-7, dev


@pytest.mark.parametrize("n", GRIDS, ids=lambda n: "x".join(map(str, n)))
def test_gas_at_rest_stays_at_rest_around_a_body(n):
    Q, geo, F = run(n, 200, quiescent=True)
    fl = geo["t"] == 0
    assert fl.sum() > 0 and (geo["phi"][fl] < 1).any()            # cut cells take part
    assert max(np.abs(Q[c][fl]).max() for c in (1, 2, 3)) < 1e-7
    assert np.abs(Q[0][fl] - 1.0).max() < 1e-7
    assert np.abs(Q[4][fl] - F[4]).max() / F[4] < 1e-7
