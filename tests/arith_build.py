"""Build recipe of tests/libarith_check.so: the step kernels' FP64 arithmetic
helpers compiled alone for gfx950 (hipcc, in-tree), for the GPU test of the
device division, reciprocal and square root.  Called by the repository's
build() next to the product library."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "arith_check.hip")
LIB = os.path.join(HERE, "libarith_check.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build() -> str:
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           SRC, "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        out = "\n".join(ln for ln in (p.stdout + p.stderr).splitlines()
                        if "packed-fp32-ops" not in ln)
        raise RuntimeError(f"hipcc failed on {SRC}:\n{out}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
