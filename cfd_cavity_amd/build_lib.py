"""Build recipe of libcavlbm.so (HIP, gfx950), in-tree.

Each translation unit is compiled by hipcc with its own floating-point
contract: the FP64 kernels, the auxiliary kernels and the host API with
``-ffp-contract=off`` (bitwise FP64 parity with an evaluation in the
reference's operand order), the FP32-arithmetic step kernels with
contraction inside each source expression only (``-ffp-contract=on``: the FMA
formation follows the source, so the single-step, multi-step and batched
kernels round alike).  Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcavlbm.so")
CLI = os.path.join(HERE, "cavity_les")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
UNITS = {
    "lbm_fp64.hip": ["-ffp-contract=off"],
    "lbm_lowp.hip": ["-ffp-contract=on"],
    "lbm_aux.hip": ["-ffp-contract=off"],
    "lbm_adjoint.hip": ["-ffp-contract=off"],
    "lbm_tangent.hip": ["-ffp-contract=off"],
    "lbm_octree.hip": ["-ffp-contract=off"],
    "cavlbm_octree.cpp": ["-ffp-contract=off", "-x", "hip"],
    "cavlbm_api.cpp": ["-ffp-contract=off", "-x", "hip"],
    "snapshot.cpp": ["-ffp-contract=off"],
}
# one source, two objects: the batched kernel in FP64 (no contraction) and in
# the FP32-arithmetic forms (contraction inside each source expression, as
# the single-cavity FP32 kernels, so a batched cavity rounds like a single one)
VARIANTS = {
    "lbm_batch_fp64": ("lbm_batch.hip", ["-ffp-contract=off", "-DCAVLBM_BATCH_PRECISION=64"]),
    "lbm_batch_lowp": ("lbm_batch.hip", ["-ffp-contract=on", "-DCAVLBM_BATCH_PRECISION=32"]),
    # the dynamic model (never contracted, by pragma) with the FP64 one-launch
    # step, and the FP32-arithmetic one-launch steps whose pull and collision
    # round like the FP32 single-step kernels
    "lbm_dynsgs": ("lbm_dynsgs.hip", ["-ffp-contract=off", "-DCAVLBM_DYN_OBJECT=64"]),
    "lbm_dynsgs_lowp": ("lbm_dynsgs.hip", ["-ffp-contract=on", "-DCAVLBM_DYN_OBJECT=32"]),
}


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    objs = []
    jobs = [(os.path.splitext(src)[0], src, extra) for src, extra in UNITS.items()]
    jobs += [(name, src, extra) for name, (src, extra) in VARIANTS.items()]
    for name, src, extra in jobs:
        obj = os.path.join(OBJ, name + ".o")
        cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        if p.returncode != 0:
            failed.append((src, "\n".join(lines)))
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the uniform-mode driver (C++ host code over the C-ABI)
    exe_tmp = CLI + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(CSRC, "cavity_les.cpp"), "-L", HERE, "-lcavlbm",
           "-Wl,-rpath,$ORIGIN", "-o", exe_tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(exe_tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
