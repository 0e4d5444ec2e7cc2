"""C ABI of libaero3d.so: every symbol of include/aero3d.h is exported, and
the library reports a missing device as an error instead of falling back."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "aero3d.h")).read() + open(os.path.join(ROOT, "include", "axi.h")).read()
    return sorted(set(re.findall(r"\b((?:aero3d|axi)_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    import aero3d
    L = aero3d.lib()
    syms = header_symbols()
    assert len(syms) >= 18
    for s in syms:
        assert hasattr(L, s), f"{s} declared in aero3d.h but not exported"
    assert set(syms) == set(aero3d.EXPORTS)


def test_create_without_device_fails_loudly():
    """Without a device the constructor raises (no CPU fallback); where the
    library finds a device the test has nothing to check."""
    import aero3d
    g = {"fx": np.arange(3.0), "fy": np.arange(3.0), "fz": np.arange(3.0)}
    N = 8
    geo = {"t": np.zeros(N, np.uint8), "phi": np.ones(N, np.float32), "ap": [np.ones(N, np.float32)] * 3 + [np.zeros(N, np.float32)]}
    try:
        s = aero3d.CfdHIP(g, geo)
    except aero3d.Aero3DError as e:
        assert "device" in str(e)
        return
    s.dispose()
    pytest.skip("a device is present")


def test_create_rejects_bad_arguments():
    import aero3d
    L = aero3d.lib()
    h = C.c_void_p()
    rc = L.aero3d_create(None, np.zeros(1, np.uint8), *[np.zeros(1, np.float32)] * 5, C.byref(h))
    assert rc == -1 and b"null" in L.aero3d_last_error(None)
    # sizes the stencil cannot take are rejected before any device work: one row
    # (the y = 0 mirror reads rows 0 and 1), an empty axis, more than 2^30 cells
    for n in ((4, 1, 3), (0, 4, 3), (1 << 11, 1 << 10, 1 << 10)):
        f = [np.arange(m + 1.0) for m in n]
        g = aero3d.Grid(*n, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in f])
        rc = L.aero3d_create(C.byref(g), np.zeros(1, np.uint8), *[np.zeros(1, np.float32)] * 5, C.byref(h))
        assert rc == -1 and b"out of range" in L.aero3d_last_error(None), n
    assert L.aero3d_destroy(None) == 0
