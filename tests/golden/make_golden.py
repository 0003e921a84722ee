"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

These are BUILD-ORACLE vectors, not reference output: the reference ships no
data files and cannot be built here (Rust toolchain and crates absent).  They
pin the oracle against regressions and give the GPU tests fixed vectors:
  oracle_16x40_210.npz  small_solver() (src/uniform/tests.rs:11-13) after 210
                        steps: f, g [k][9]; rho, ux, uy, T, nu_t, fy_node.
  digests.json          sha256 of the FP64 state bytes after N steps for the
                        64x256 Ra=1e6 and 128x512 Ra=1e8 configurations, plus
                        the diagnostics at that step.
Run: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

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


def state_digest(s) -> str:
    h = hashlib.sha256()
    for a in (s.f, s.g, *(getattr(s, n) for n in FIELDS)):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def main():
    s = pyoracle.OracleSolver(16, 40, 1.0e5, 0.71, 0.1, 0.9, 0.02, 42, 1)
    s.run(210)
    np.savez_compressed(os.path.join(HERE, "oracle_16x40_210.npz"), f=s.f, g=s.g,
                        **{n: getattr(s, n) for n in FIELDS})
    digests = {}
    for name, (args, steps) in CASES.items():
        o = pyoracle.OracleSolver(*args, 8)
        o.run(steps)
        d = o.diagnostics()
        digests[name] = {"args": list(args), "steps": steps, "sha256": state_digest(o),
                         "diagnostics": {k: v for k, v in d.items()}}
    with open(os.path.join(HERE, "digests.json"), "w") as fh:
        json.dump(digests, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
