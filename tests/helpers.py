"""Shared helpers for the parity tests."""
import numpy as np

SMALL = dict(nx=16, ny=40, ra=1.0e5, pr=0.71, cs=0.1, pr_t=0.9, utarget=0.02, seed=42)


def args_of(cfg):
    return (cfg["nx"], cfg["ny"], cfg["ra"], cfg["pr"], cfg["cs"], cfg["pr_t"],
            cfg["utarget"], cfg["seed"])


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def first_mismatch(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    bad = np.nonzero(a.view(np.uint64) != b.view(np.uint64))[0]
    if bad.size == 0:
        return "equal"
    k = bad[0]
    return f"{bad.size} mismatches; first at {k}: {a[k]!r} vs {b[k]!r}"


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)
