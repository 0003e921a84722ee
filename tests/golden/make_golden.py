"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

These are BUILD-ORACLE vectors, not reference output: the reference ships no
data files and cannot be built here (Rust toolchain and crates absent).  They
pin the oracle against regressions and give the GPU tests fixed vectors:
  oracle_16x40_210.npz  small_solver() (src/uniform/tests.rs:11-13) after 210
                        steps: f, g [k][9]; rho, ux, uy, T, nu_t, fy_node.
  digests.json          sha256 of the FP64 state bytes after N steps, plus the
                        diagnostics at that step, for
                        - 64x256 Ra=1e6 and 128x512 Ra=1e8 after 1000 steps;
                        - 64x256 Ra=1e6 after 20 000 steps with the
                          diagnostics logged every 1000 steps (the
                          reference's own CPU configuration);
                        - 128x512 Ra=1e8 after 100 000 steps (the GPU drop-in
                          configuration's last checkpoint);
                        - 256x1024 Ra=1e10, utarget 0.005, after 10 000 steps
                          (the precision-form configuration in FP64);
                        - each of the 64 seeded batch cavities
                          (cfd_cavity_amd.batch.seeded_batch_specs) after 300
                          steps.
Run: python tests/golden/make_golden.py [--long] [--batch]
(the default writes the short cases; --long and --batch add the cases that
take minutes of oracle time and keep the others as they are).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402

FIELDS = ("rho", "ux", "uy", "temperature", "nu_t", "fy_node")
CASES = {
    "64x256_ra1e6_1000": ((64, 256, 1.0e6, 0.71, 0.1, 0.9, 0.02, 42), 1000),
    "128x512_ra1e8_1000": ((128, 512, 1.0e8, 0.71, 0.1, 0.9, 0.02, 42), 1000),
}
LONG_CASES = {
    "128x512_ra1e8_100000": ((128, 512, 1.0e8, 0.71, 0.1, 0.9, 0.02, 42), 100000),
    "256x1024_ra1e10_10000": ((256, 1024, 1.0e10, 0.71, 0.1, 0.9, 0.005, 42), 10000),
}
# name: (args, steps, log interval)
LOG_CASES = {
    "64x256_ra1e6_20000_log1000": ((64, 256, 1.0e6, 0.71, 0.1, 0.9, 0.02, 42), 20000, 1000),
}
BATCH_STEPS = 300
BATCH_KEY = "batch64_fp64_300"
JOBS = int(os.environ.get("GOLDEN_JOBS", os.cpu_count() or 8))


def state_digest(s) -> str:
    h = hashlib.sha256()
    for a in (s.f, s.g, *(getattr(s, n) for n in FIELDS)):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def spec_args(sp):
    return (sp["nx"], sp["ny"], sp["ra"], sp["pr"], sp["cs"], sp["pr_t"], sp["utarget"],
            sp["seed"])


def run_case(args, steps):
    t0 = time.time()
    o = pyoracle.OracleSolver(*args, JOBS)
    o.run(steps)
    d = o.diagnostics()
    print(f"  {args[0]}x{args[1]} x {steps}: {time.time() - t0:.1f} s", flush=True)
    return {"args": list(args), "steps": steps, "sha256": state_digest(o),
            "diagnostics": {k: v for k, v in d.items()}}


def run_log_case(args, steps, every):
    t0 = time.time()
    o = pyoracle.OracleSolver(*args, JOBS)
    log = []
    while o.time_step < steps:
        o.run(every)
        log.append({"step": o.time_step, **o.diagnostics()})
    print(f"  {args[0]}x{args[1]} x {steps}, log every {every}: {time.time() - t0:.1f} s",
          flush=True)
    return {"args": list(args), "steps": steps, "every": every, "sha256": state_digest(o),
            "log": log}


def main(argv):
    path = os.path.join(HERE, "digests.json")
    digests = {}
    if os.path.exists(path):
        with open(path) as fh:
            digests = json.load(fh)
    s = pyoracle.OracleSolver(16, 40, 1.0e5, 0.71, 0.1, 0.9, 0.02, 42, 1)
    s.run(210)
    np.savez_compressed(os.path.join(HERE, "oracle_16x40_210.npz"), f=s.f, g=s.g,
                        **{n: getattr(s, n) for n in FIELDS})
    todo = dict(CASES)
    if "--long" in argv:
        todo.update(LONG_CASES)
    for name, (args, steps) in todo.items():
        digests[name] = run_case(args, steps)
    if "--long" in argv:
        for name, (args, steps, every) in LOG_CASES.items():
            digests[name] = run_log_case(args, steps, every)
    if "--batch" in argv:
        from cfd_cavity_amd.batch import seeded_batch_specs

        digests[BATCH_KEY] = [run_case(spec_args(sp), BATCH_STEPS) for sp in seeded_batch_specs()]
    with open(path, "w") as fh:
        json.dump(digests, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
