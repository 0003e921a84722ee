"""The step kernels' FP64 division, reciprocal and square root on the GPU.

The FP64 kernels do not use the compiler's division and square-root
sequences: quotients by the constant divisors take one Markstein correction
(Arith::divc), quotients by densities, Pr_t and the dynamic model's window
counts take a correctly rounded reciprocal and two corrections (Arith::divr,
DynWin::mean), and the remaining divisions, reciprocals and |S|'s square root
use the hardware's first approximation refined by FMA steps without operand
scaling (Arith::div, Arith::rcp, Arith::sqrtp).  Bitwise parity with the
oracle rests on each of them returning the IEEE result.  This runs the
helpers themselves (tests/arith_check.hip includes lbm_device.hpp) on the
GPU and compares them bit for bit with the host's IEEE division and sqrt on:
random operands over 2^-60..2^60, the ranges the kernels meet, divisors
whose significand is all ones (where the reciprocal iteration has its one
known failure unless its first approximation is already correctly rounded),
quotients A/B within 1/(2B) ulp of a rounding midpoint (2^54 A = (2Q+1) B +- 1
for significands A < B), the constant divisors, the window counts 1..64,
exact square roots, roots within about 2^-96 relative of a rounding midpoint
(X 2^55 = s^2 - t for small t of either sign), +0 and infinity, and random
square roots from 2^-767 to the largest double.  The CPU test of the same sequences (test_division_algorithm.py)
assumes a first approximation within 2^-20; the hardware's is checked here
to be well inside that.
"""
import ctypes
import math
import os
import random

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libarith_check.so")
CS2 = 1.0 / 3.0
CONST_DIVISORS = (CS2, 2.0 * CS2 * CS2, 2.0 * CS2, CS2 * CS2, 3.0)


def load():
    if not os.path.exists(LIB):
        raise RuntimeError(f"{LIB} is missing: run __graft_entry__.build() first")
    lib = ctypes.CDLL(LIB)
    lib.arith_check.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int]
    lib.arith_check.restype = ctypes.c_int
    lib.arith_check_planes.restype = ctypes.c_int
    return lib


def near_midpoint_pairs(count, seed):
    """(A, B) 53-bit significands, A < B, with 2^54 A = (2Q+1) B +- 1: A/B
    lies in (1/2, 1), where doubles are multiples of 2^-53, within 2^-54/B of
    the midpoint (2Q+1) 2^-54 between two of them."""
    r = random.Random(seed)
    out = []
    b1 = (1 << 53) - 1
    for s in (1, -1):
        a1 = (s * pow(1 << 54, -1, b1)) % b1
        if a1 >= 1 << 52:
            out.append((a1, b1))
    while len(out) < count:
        b = r.randrange(1 << 52, 1 << 53) | 1
        a = (r.choice((1, -1)) * pow(1 << 54, -1, b)) % b
        if a >= 1 << 52:
            out.append((a, b))
    return out


def sqrt_mod_pow2(t, k):
    """r with r*r = t (mod 2^k) for t = 1 (mod 8), lifted bit by bit: from
    r*r = t (mod 2^i), adding 2^(i-1) flips bit i of r*r (i >= 3)."""
    m = 1 << k
    t %= m
    r = 1
    for i in range(3, k):
        if ((r * r - t) >> i) & 1:
            r += 1 << (i - 1)
    return r % m


def near_midpoint_roots(count, sign, t_start=1):
    """53-bit significands X with X 2^55 = s^2 - t for an odd s in
    [2^53.5, 2^54) and a small t = 1 (mod 8) of the given sign: sqrt(X 2^-53)
    lies in [2^-0.5, 1) within about |t| 2^-109 of the midpoint s 2^-54 between
    two doubles, below it for t > 0 and above it for t < 0.  Returns
    (X, t) pairs, |t| growing from t_start."""
    out = []
    t = t_start if sign > 0 else -(t_start + 6)
    while len(out) < count:
        r = sqrt_mod_pow2(t, 55)
        m = 1 << 55
        for s in sorted({r, m - r, (r + (m >> 1)) % m, (m - r + (m >> 1)) % m}):
            if s % 2 == 1 and s * s >= 1 << 107 and s < 1 << 54:
                x = (s * s - t) >> 55
                if (1 << 52) <= x < (1 << 53) and (s * s - t) % m == 0:
                    out.append((x, t))
        t += 8 * sign
    return out[:count]


def operand_groups():
    rng = np.random.default_rng(20240611)
    sign = lambda n: rng.choice([-1.0, 1.0], n)  # noqa: E731
    groups = {}
    n = 1 << 19
    groups["random"] = (np.exp2(rng.uniform(-60, 60, n)) * sign(n),
                        np.exp2(rng.uniform(-60, 60, n)) * sign(n),
                        np.exp2(rng.uniform(-140, 60, n)))
    n = 1 << 18
    groups["kernel ranges"] = (rng.uniform(-1.0, 1.0, n), rng.uniform(0.5, 2.0, n),
                               np.exp2(rng.uniform(-125, -4, n)))
    ones = np.array([np.ldexp(2.0 - 2.0 ** -52, e) for e in range(-60, 61)])
    ones = np.concatenate([ones, -ones])
    m = len(ones)
    groups["all-ones divisors"] = (np.concatenate([np.ones(m), rng.uniform(0.5, 2.0, m) * sign(m)]),
                                   np.concatenate([ones, ones]),
                                   np.concatenate([np.abs(ones), np.abs(ones) * 0.5]))
    pairs = near_midpoint_pairs(20000, 7)
    r = random.Random(8)
    scale = [r.randrange(-20, 20) for _ in pairs]
    groups["near-midpoint quotients"] = (
        np.array([np.ldexp(float(a), e - 52) for (a, _), e in zip(pairs, scale)]),
        np.array([np.ldexp(float(b), e - 52 - r.randrange(-20, 20)) for (_, b), e in zip(pairs, scale)]),
        np.exp2(rng.uniform(-60, 60, len(pairs))))
    n = 1 << 16
    ints = rng.integers(1, 1 << 26, n).astype(np.float64)
    near1 = 1.0 + rng.integers(-4096, 4096, n) * 2.0 ** -52
    r = random.Random(9)
    hard = [math.ldexp(float(x), r.randrange(-30, 30) * 2 - 53)
            for sg in (1, -1) for x, _ in near_midpoint_roots(6000, sg)]
    xs = np.concatenate([ints * ints, near1, hard, [0.0, np.inf]])
    groups["square roots"] = (np.ones(len(xs)), np.ones(len(xs)), xs)
    # the dynamic model's square roots rely on sqrtp from 2^-767 (where the
    # generic sequence starts to scale) to the largest double
    n = 1 << 17
    xw = np.concatenate([np.exp2(rng.uniform(-767, 1023.99, n - 4)),
                         [2.0 ** -767, np.nextafter(2.0 ** -767, 1.0),
                          np.finfo(np.float64).max, np.nextafter(np.finfo(np.float64).max, 0.0)]])
    groups["square roots, full range"] = (np.ones(n), np.ones(n), xw)
    n = 1 << 17
    groups["window counts"] = (rng.uniform(-64.0, 64.0, n) * np.exp2(rng.uniform(-50, 0, n)),
                               rng.integers(1, 65, n).astype(np.float64),
                               np.ones(n))
    groups["constant divisors"] = (np.exp2(rng.uniform(-60, 4, n)) * sign(n),
                                   rng.choice(np.array(CONST_DIVISORS), n),
                                   np.ones(n))
    return groups


def run_device(lib, a, b, x, y):
    n = len(a)
    ops = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, x, y)]
    assert all(v.shape == (n,) for v in ops)
    planes = lib.arith_check_planes()
    out = np.empty(planes * n, dtype=np.float64)
    rc = lib.arith_check(*(v.ctypes.data for v in ops), out.ctypes.data, n)
    assert rc == 0, f"arith_check returned {rc}"
    return out.reshape(planes, n)


def mismatches(got, want, a, b, x):
    bad = np.nonzero(got.view(np.uint64) != want.view(np.uint64))[0]
    return len(bad), [(a[j].hex(), b[j].hex(), x[j].hex(), got[j].hex(), want[j].hex())
                      for j in bad[:4]]


@pytest.mark.gpu
@pytest.mark.parametrize("group", list(operand_groups().keys()))
def test_device_helpers_equal_ieee(group):
    lib = load()
    a, b, x = operand_groups()[group]
    y = 1.0 / b
    o = run_device(lib, a, b, x, y)
    quot = a / b
    checks = {"Arith::rcp": (o[2], y), "Arith::sqrtp": (o[4], np.sqrt(x)),
              "DynWin::mean": (o[7], quot), "Arith::divr": (o[5], quot)}
    if not group.startswith("square roots"):
        checks["Arith::div"] = (o[3], quot)
    if group == "constant divisors":
        checks["Arith::divc"] = (o[6], quot)
    for name, (got, want) in checks.items():
        nbad, first = mismatches(got, want, a, b, x)
        assert nbad == 0, f"{group}: {name} differs from IEEE in {nbad} cases, e.g. (a, b, x, got, want) {first}"
    # The hardware's first approximations, on finite nonzero operands, lie
    # well inside the 2^-20 the CPU test of the refinement assumes.
    fin = np.isfinite(x) & (x > 0)
    e_rcp = np.max(np.abs(o[0] * b - 1.0))
    e_rsq = np.max(np.abs(o[1][fin] * np.sqrt(x[fin]) - 1.0))
    assert e_rcp < 2.0 ** -22 and e_rsq < 2.0 ** -22, (e_rcp, e_rsq)


def test_operands_distinguish_the_corrections():
    """Plain multiplication by the reciprocal, the mistake the corrections
    exist to fix, differs from IEEE division on the constant divisors and on
    the near-midpoint quotients (host arithmetic), so the checks above have
    teeth for divc, divr and DynWin::mean."""
    g = operand_groups()
    for name in ("constant divisors", "near-midpoint quotients"):
        a, b, _ = g[name]
        assert np.count_nonzero(a * (1.0 / b) != a / b) > len(a) // 10, name


# ---------------------------------------------------------------------------
# Stochastic rounding of the FP16-offset storage (sr_half: FP32 -> FP16 with
# v_cvt_sr_f16_f32).  The storage relies on three properties: the result is
# one of the two FP16 neighbours of x; it is the neighbour away from zero
# exactly when the top 13 bits of the random word plus the 13 significand
# bits that the conversion drops carry out (FP16 normal results); and so its
# mean over the 8192 values of those 13 bits is x (exactly for normal
# results, within 2^-37 where the result is subnormal).

FP16_MIN_NORMAL = 2.0 ** -14


def sr_values():
    rng = np.random.default_rng(20261204)
    sign = lambda n: rng.choice([-1.0, 1.0], n)  # noqa: E731
    parts = [
        np.exp2(rng.uniform(-14, 15, 1536)) * sign(1536),     # FP16 normal range
        np.exp2(rng.uniform(-16, -2, 1536)) * sign(1536),     # the stored deviations' range
        np.exp2(rng.uniform(-30, -14, 768)) * sign(768),      # FP16 subnormal range
        rng.integers(-2048, 2048, 254) * 2.0 ** -12,          # exact FP16 values
        [0.0, -0.0],
    ]
    return np.concatenate([np.asarray(p, dtype=np.float64) for p in parts]).astype(np.float32)


def sr_words():
    rng = np.random.default_rng(77)
    top = np.arange(8192, dtype=np.uint64) << np.uint64(19)   # every top-13-bit value, low bits 0
    rand = rng.integers(0, 1 << 32, 2048, dtype=np.uint64)
    return np.concatenate([top, rand]).astype(np.uint32)


def fp16_neighbours(x):
    """(toward zero, away from zero) FP16 neighbours of the FP32 values x, as
    float64; equal where x is an FP16 value."""
    xd = x.astype(np.float64)
    h = x.astype(np.float16)
    hd = h.astype(np.float64)
    down = np.where(hd <= xd, h, np.nextafter(h, np.float16(-np.inf))).astype(np.float64)
    up = np.where(hd >= xd, h, np.nextafter(h, np.float16(np.inf))).astype(np.float64)
    toward = np.where(xd >= 0, down, up)
    away = np.where(xd >= 0, up, down)
    return toward, away


def test_sr_operands_distinguish_rounding_to_nearest():
    """Rounding to nearest, the rounding the storage replaced, fails the mean
    check on most of the values above (host arithmetic), so the check has
    teeth."""
    x = sr_values()
    rn = x.astype(np.float16).astype(np.float64)
    assert np.count_nonzero(rn != x.astype(np.float64)) > len(x) // 2


@pytest.mark.gpu
def test_stochastic_rounding_to_fp16():
    lib = load()
    lib.sr_check.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int] * 2
    lib.sr_check.restype = ctypes.c_int
    x = np.ascontiguousarray(sr_values())
    r = np.ascontiguousarray(sr_words())
    nx, nr = len(x), len(r)
    out = np.empty((nr, nx), dtype=np.uint16)
    rc = lib.sr_check(x.ctypes.data, r.ctypes.data, out.ctypes.data, nx, nr)
    assert rc == 0, f"sr_check returned {rc}"
    got = out.view(np.float16).astype(np.float64)
    xd = x.astype(np.float64)
    toward, away = fp16_neighbours(x)
    # one of the two neighbours, for every word
    ok = (got == toward[None, :]) | (got == away[None, :])
    assert ok.all(), f"{np.count_nonzero(~ok)} results outside the neighbours, e.g. x " \
        f"{xd[np.nonzero(~ok)[1][:4]]}"
    # the carry rule, on normal results, for every word
    normal = np.abs(toward) >= FP16_MIN_NORMAL
    dropped = (x.view(np.uint32) & np.uint32(0x1FFF)).astype(np.int64)
    carry = dropped[None, :] + (r.astype(np.int64) >> 19)[:, None] >= 8192
    want = np.where(carry, away[None, :], toward[None, :])
    bad = (got != want) & normal[None, :]
    assert not bad.any(), f"{np.count_nonzero(bad)} normal results break the carry rule"
    # the mean over every value of the top 13 bits
    mean = got[:8192].sum(axis=0) / 8192.0
    assert np.array_equal(mean[normal], xd[normal])
    err = np.abs(mean[~normal] - xd[~normal])
    assert err.max() <= 2.0 ** -37, err.max()
