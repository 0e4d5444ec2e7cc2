"""Seeded synthetic cases for grids of any size, down to one cell per axis.

An ellipsoid on the symmetry plane y = 0 cut by a mildly stretched grid: volume
fractions from 27 sub-points per cell and + face apertures from 9 sub-points
per face (the reference's sampling, src/air_cfd.js:68-90), flag 1 below a
fraction of 0.2, faces next to a solid cell closed, a few partial z faces
with baffle slopes, and one cell of each fixed-state flag (3, 4, 5) where the
grid has room. The masks follow the reference's construction rules, not its
body shapes, so they exercise every branch of the stencil (cut, solid, mirror,
freestream ghosts, baffle, fixed cells) on grids the reference never builds.
"""
import math

import numpy as np

SUB = np.array([1 / 6, 0.5, 5 / 6])


def _faces(rng, n, lo):
    return np.concatenate([[0.0], np.cumsum(1.0 + 0.15 * rng.random(n))]) + lo


def synth_case(NX, NY, NZ, seed=0, body=True, quiescent=False, mach=0.6):
    """-> (grid, geo, F, fix) in the layouts the solver, the oracle and the batch take.
    body=False: no body, no baffles, no fixed cells (every cell fluid, every face
    open). quiescent=True: gas at rest (M = 0) with fixed states equal to it.
    mach: freestream Mach number along +x (negative: towards -x); |mach| > 1
    makes most faces supersonic (the HLLC's upwind-side branches)."""
    rng = np.random.default_rng(seed)
    fx, fy, fz = _faces(rng, NX, -0.5 * NX), _faces(rng, NY, 0.0), _faces(rng, NZ, -0.5 * NZ)
    cx, cz = 0.5 * (fx[0] + fx[-1]), 0.5 * (fz[0] + fz[-1])
    ax, ay, az = 0.3 * (fx[-1] - fx[0]) + 0.3, 0.55 * fy[-1] + 0.3, 0.3 * (fz[-1] - fz[0]) + 0.3

    def solid(x, y, z):
        if not body:
            return np.zeros(np.broadcast(x, y, z).shape, bool)
        return ((x - cx) / ax) ** 2 + (y / ay) ** 2 + ((z - cz) / az) ** 2 < 1.0

    dx, dy, dz = np.diff(fx), np.diff(fy), np.diff(fz)
    X = fx[:-1, None] + SUB[None, :] * dx[:, None]          # (NX, 3) sub-point coordinates
    Y = fy[:-1, None] + SUB[None, :] * dy[:, None]
    Z = fz[:-1, None] + SUB[None, :] * dz[:, None]
    # volume fraction: 27 sub-points per cell, array index [k, j, i]
    s = solid(X[None, None, :, None, None, :], Y[None, :, None, None, :, None], Z[:, None, None, :, None, None])
    phi = (27 - s.reshape(NZ, NY, NX, 27).sum(-1)) / 27.0
    ap = []
    # + x face at fx[i+1]: sub-points over (y, z); + y face over (x, z); + z face over (x, y)
    sxf = solid(fx[1:][None, None, :, None, None], Y[None, :, None, :, None], Z[:, None, None, None, :])
    ap.append((9 - sxf.reshape(NZ, NY, NX, 9).sum(-1)) / 9.0)
    syf = solid(X[None, None, :, :, None], fy[1:][None, :, None, None, None], Z[:, None, None, None, :])
    ap.append((9 - syf.reshape(NZ, NY, NX, 9).sum(-1)) / 9.0)
    szf = solid(X[None, None, :, :, None], Y[None, :, None, None, :], fz[1:][:, None, None, None, None])
    ap.append((9 - szf.reshape(NZ, NY, NX, 9).sum(-1)) / 9.0)
    t = (phi < 0.2).astype(np.uint8)
    # faces next to a solid cell are closed (src/air_cfd.js:132-139)
    sol = t == 1
    ap[0][:, :, :-1][sol[:, :, :-1] | sol[:, :, 1:]] = 0
    ap[1][:, :-1, :][sol[:, :-1, :] | sol[:, 1:, :]] = 0
    ap[2][:-1][sol[:-1] | sol[1:]] = 0
    ap3 = np.zeros((NZ, NY, NX))
    fluid = np.argwhere((t == 0) & (np.arange(NZ)[:, None, None] + 1 < NZ))
    if len(fluid) and body:
        pick = fluid[rng.random(len(fluid)) < 0.05]
        for k, j, i in pick:
            ap[2][k, j, i] = min(ap[2][k, j, i], 1 - 3 / 9)
            ap3[k, j, i] = rng.uniform(-0.12, 0.12)
    full = np.argwhere((t == 0) & (phi == 1.0))
    if len(full) >= 6 and body:
        for flag, idx in zip((3, 4, 5), rng.choice(len(full), 3, replace=False)):
            k, j, i = full[idx]
            t[k, j, i] = flag
    # no baffle slope on a z face of or next to a fixed cell, none on the top boundary
    fixed = t >= 3
    ap3[fixed] = 0
    ap3[:-1][fixed[1:]] = 0
    ap3[-1] = 0
    grid = {"NX": NX, "NY": NY, "NZ": NZ, "fx": fx, "fy": fy, "fz": fz}
    for a, f in (("x", fx), ("y", fy), ("z", fz)):
        grid["d" + a] = (f[1:] - f[:-1]).astype(np.float32)
        grid[a + "c"] = (0.5 * (f[:-1] + f[1:])).astype(np.float32)
    geo = {"t": t.reshape(-1), "phi": phi.reshape(-1).astype(np.float32),
           "ap": [a.reshape(-1).astype(np.float32) for a in ap] + [ap3.reshape(-1).astype(np.float32)]}
    M, al = (0.0 if quiescent else mach), math.radians(4.0)
    p0 = 1 / 1.4
    F = [1.0, M * math.cos(al), 0.0, M * math.sin(al), p0 / 0.4 + 0.5 * M * M]

    def state(r, u, p):
        return [r, r * u * M, 0.0, 0.0, p * p0 / 0.4 + 0.5 * r * (u * M) ** 2]

    fix = [list(F)] * 3 if quiescent else [state(1.05, 0.45, 1.1), state(0.95, 1.0, 1.3), state(0.5, 1.6, 1.2)]
    return grid, geo, F, fix


SIZES = [(1, 2, 1), (1, 2, 2), (2, 2, 3), (3, 5, 2), (5, 3, 7), (63, 7, 5), (64, 8, 4), (65, 9, 3),
         (130, 17, 6), (7, 40, 2), (200, 2, 1)]
