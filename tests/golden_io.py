"""Loader of the golden vectors under tests/golden (written by make_golden.js)."""
import gzip
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_DT = {"f4": np.float32, "f8": np.float64, "u1": np.uint8, "i1": np.int8, "i4": np.int32}


class Golden:
    def __init__(self, case):
        self.dir = os.path.join(GOLDEN, case)
        with open(os.path.join(self.dir, "manifest.json")) as f:
            self.m = json.load(f)
        self.s = self.m["scalars"]

    def has(self, name):
        a = self.m["arrays"].get(name)
        return a is not None and a["file"] is not None

    def arr(self, name):
        a = self.m["arrays"][name]
        with gzip.open(os.path.join(self.dir, a["file"])) as f:
            raw = f.read()
        assert hashlib.sha256(raw).hexdigest() == a["sha256"], f"fixture {name} corrupted"
        return np.frombuffer(raw, dtype=_DT[a["dtype"]]).copy()

    def sha(self, name):
        return self.m["arrays"][name]["sha256"]

    def grid(self):
        fx, fy, fz = self.arr("fx"), self.arr("fy"), self.arr("fz")
        g = {"NX": len(fx) - 1, "NY": len(fy) - 1, "NZ": len(fz) - 1, "fx": fx, "fy": fy, "fz": fz}
        # Float32 widths and centres computed from Float64 faces, as the reference grid does
        for ax, f in (("x", fx), ("y", fy), ("z", fz)):
            g["d" + ax] = (f[1:] - f[:-1]).astype(np.float32)
            g[ax + "c"] = (0.5 * (f[:-1] + f[1:])).astype(np.float32)
        return g

    def geo(self):
        return {"t": self.arr("t"), "phi": self.arr("phi"), "ap": [self.arr(f"ap{d}") for d in range(4)]}

    def force_list(self):
        return {k: self.arr("fl_" + k) for k in ("q", "cx", "cz", "x", "z")}

    def state(self, tag):
        return np.stack([self.arr(f"{tag}_{v}") for v in ("r", "ru", "rv", "rw", "E")])


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
