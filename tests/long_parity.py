"""Long bitwise comparison of the GPU path with the oracle (not collected by
pytest; run it on a GPU machine):

    python -m tests.long_parity --nx 128 --ny 512 --ra 1e8 --steps 1000,10000,100000

The 128x512, Ra=1e8 case is the GPU drop-in measurement configuration: FP64
with the default steps per launch, compared at each listed time step for the
populations, the fields, fy_node and nu_t (bitwise) and the diagnostics
(1e-12 relative; maxima exact).  Prints one JSON line per checkpoint and
exits non-zero at the first difference.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from cfd_cavity_amd import CavitySolver  # noqa: E402
from oracle import pyoracle  # noqa: E402
from tests.helpers import SMALL, args_of, bits_equal, first_mismatch, rel  # noqa: E402

FIELDS = ("rho", "ux", "uy", "temperature", "nu_t", "fy_node")


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--ra", type=float, default=1.0e8)
    ap.add_argument("--steps", default="1000,10000,100000")
    ap.add_argument("--jobs", type=int, default=min(os.cpu_count() or 1, 16))
    a = ap.parse_args()
    cfg = dict(SMALL, nx=a.nx, ny=a.ny, ra=a.ra)
    g = CavitySolver(*args_of(cfg))
    o = pyoracle.OracleSolver(*args_of(cfg), a.jobs)
    ok = True
    for upto in (int(x) for x in a.steps.split(",")):
        t0 = time.perf_counter()
        g.run(upto - g.time_step)
        g.synchronize()
        tg = time.perf_counter() - t0
        t0 = time.perf_counter()
        o.run(upto - o.time_step)
        to = time.perf_counter() - t0
        bad = []
        for name in FIELDS:
            x, y = g.get_field(name), getattr(o, name)
            if not bits_equal(x, y):
                bad.append(f"{name}: {first_mismatch(x, y)}")
        for w in ("f", "g"):
            x, y = g.get_pops(w), getattr(o, w)
            if not bits_equal(x, y):
                bad.append(f"pops {w}: {first_mismatch(x, y)}")
        dg, do = g.diagnostics(), o.diagnostics()
        drel = max(rel(dg[k], do[k]) for k in ("nusselt_hot", "nusselt_cold", "total_mass"))
        print(json.dumps({"time_step": upto, "bitwise": not bad, "mismatches": bad,
                          "nusselt_hot": dg["nusselt_hot"], "total_mass": dg["total_mass"],
                          "diag_rel_max": drel, "steps_per_launch": g.steps_per_launch,
                          "gpu_s": round(tg, 3), "oracle_s": round(to, 1),
                          "oracle_threads": a.jobs}), flush=True)
        if bad or drel > 1e-12:
            ok = False
            break
    sys.exit(0 if ok else 1)


if __name__ == "__main__":
    main()
