"""CPU build of the engine field model's device functions against the
reference's own outputs:
    python tests/cpu_emul/run_field_emul.py <lib>   -> one JSON line
(samples at the seeded points, initial states of the setups, the table check
on corrupted tables)."""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "3d_aero_lab_amd"))
import aero3d  # noqa: E402  (table builder only; no device use)
import field_cases as FC  # noqa: E402


def main():
    L = C.CDLL(sys.argv[1])
    P = lambda a: a.ctypes.data_as(C.c_void_p)

    def sample(desc, x, r, th):
        T = aero3d.field_table(desc)
        out = np.empty(14 * x.size)
        L.fm_sample(P(T), P(np.ascontiguousarray(x)), P(np.ascontiguousarray(r)), P(np.ascontiguousarray(th)), int(x.size), P(out))
        return out.reshape(-1, 14)

    def init(desc, xf, rf, solid, stream, uniform):
        T = aero3d.field_table(desc)
        W, H = xf.size - 1, rf.size - 1
        Q = np.empty(5 * W * H)
        L.fm_initial(P(T), W, H, P(np.ascontiguousarray(xf)), P(np.ascontiguousarray(rf)), P(np.ascontiguousarray(solid, np.uint8)),
                     P(np.ascontiguousarray(stream, np.int8)), int(uniform), P(Q))
        return Q.reshape(5, -1)

    # corrupted tables: each must be refused before any evaluation
    from golden_io import Golden
    T0 = aero3d.field_table(Golden("axi_field").s["models"]["hbtf"])
    why = C.create_string_buffer(256)
    bad = {}
    for name, (i, v) in {"version": (0, 2.0), "type": (1, 7.0), "nstream": (3, 5.0), "wall_offset": (7, float(T0.size)),
                         "knot_count": (12, 1e6), "body_count": (44, 17.0), "body_offset": (49, -4.0), "comb_stream": (141, 3.0),
                         "liner_points": (149, 0.0), "outer_offset": (177, 0.5), "exits": (182, 5.0), "nan_count": (4 + 10 + 4, float("nan"))}.items():
        T = T0.copy()
        T[i] = v
        bad[name] = L.fm_check(P(T), int(T.size), why, 256) != 0
    bad["short"] = L.fm_check(P(T0), 100, why, 256) != 0
    ok = L.fm_check(P(T0), int(T0.size), why, 256) == 0
    print(json.dumps({"samples": FC.compare_samples(sample), "initial": FC.compare_initial_states(init),
                      "rejected": bad, "sound_table_accepted": ok}))


if __name__ == "__main__":
    main()
