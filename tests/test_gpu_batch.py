"""Batched form.  In FP64 every cavity of a batch is bitwise the oracle's
cavity: the seeded 64-cavity workload (sizes 64x256 .. 256x1024, Ra in
[1e6, 1e9], Cs in [0.08, 0.12]) against the oracle's per-cavity digests after
300 steps, and a small mixed batch against live oracle runs.  In every storage
a batched cavity equals, bit for bit, the same cavity stepped alone (same
per-node code), across graph replays and after the autotuner (which must
leave the state unchanged); the FP32 workload stays finite and bounded for
its 50 000 steps.  A batch advances several steps per launch (the N-step
tiles) like a single cavity does; the comparisons run each side with its own
choice, so they also check the batched N-step kernel against the
single-cavity kernels."""
import os

import numpy as np
import pytest

from cfd_cavity_amd import CavityBatch, CavitySolver
from cfd_cavity_amd.batch import seeded_batch_specs
from tests.golden.make_golden import BATCH_KEY, BATCH_STEPS, FIELDS
from tests.helpers import bits_equal
from tests.test_golden import assert_diagnostics_close, device_digest, load_digests

pytestmark = pytest.mark.gpu

SPECS = [
    dict(nx=16, ny=40, ra=1e5, cs=0.10),
    dict(nx=64, ny=256, ra=1e6, cs=0.08),
    dict(nx=96, ny=384, ra=3e7, cs=0.12),
    dict(nx=24, ny=100, ra=2e6, cs=0.11, seed=7),
]


def single(sp, storage):
    return CavitySolver(sp["nx"], sp["ny"], sp["ra"], 0.71, sp["cs"], 0.9, 0.02,
                        sp.get("seed", 42), 1, storage=storage)


def same_state(a, b):
    for name in ("rho", "ux", "uy", "temperature", "fy_node", "nu_t"):
        if not bits_equal(a.get_field(name), b.get_field(name)):
            return name
    for w in ("f", "g"):
        if not bits_equal(a.get_pops(w), b.get_pops(w)):
            return w
    return None


def test_seeded_batch64_fp64_matches_oracle_digests():
    """The 64 seeded cavities in one FP64 batch, run in chunks that cover the
    spin-up's direct launches, graph replays and a remainder (1 + 150 + 149
    steps), are each bitwise the oracle's cavity after 300 steps."""
    specs = seeded_batch_specs()
    want = load_digests()[BATCH_KEY]
    assert len(want) == len(specs) == 64
    b = CavityBatch(specs, storage="fp64")
    for chunk in (1, 150, BATCH_STEPS - 151):
        b.run(chunk)
    for i, (c, sp, dig) in enumerate(zip(b, specs, want)):
        assert dig["args"][:2] == [sp["nx"], sp["ny"]] and dig["steps"] == BATCH_STEPS
        assert c.time_step == BATCH_STEPS
        assert device_digest(c) == dig["sha256"], f"cavity {i} ({sp['nx']}x{sp['ny']})"
        assert_diagnostics_close(c.diagnostics(), dig["diagnostics"])


def test_small_batch_fp64_matches_live_oracle(oracle):
    """A mixed FP64 batch (one lattice narrower than a tile) against the
    oracle run beside it, after the spin-up and after graph replays."""
    batch = CavityBatch(SPECS, storage="fp64")
    jobs = min(16, os.cpu_count() or 1)
    ors = [oracle.OracleSolver(sp["nx"], sp["ny"], sp["ra"], 0.71, sp["cs"], 0.9, 0.02,
                               sp.get("seed", 42), jobs) for sp in SPECS]
    for chunk in (3, 250):
        batch.run(chunk)
        for o in ors:
            o.run(chunk)
        for i, (c, o) in enumerate(zip(batch, ors)):
            assert bits_equal(c.get_pops("f"), o.f) and bits_equal(c.get_pops("g"), o.g), i
            for n in FIELDS:
                assert bits_equal(c.get_field(n), getattr(o, n)), (i, n)


def test_seeded_batch64_fp32_50k_steps_stays_bounded():
    """The batched workload as specified: 64 seeded cavities, FP32 storage,
    50 000 steps.  Every cavity stays finite, under the reference's abort
    threshold max|u| < 0.10 (src/main.rs:215-275), with its mass within 1e-4
    of the initial nx*ny."""
    specs = seeded_batch_specs()
    b = CavityBatch(specs, storage="fp32")
    b.run(50000)
    worst = (0.0, 0.0)
    for i, (c, sp) in enumerate(zip(b, specs)):
        d = c.diagnostics()
        n = sp["nx"] * sp["ny"]
        drift = abs(d["total_mass"] - n) / n
        worst = (max(worst[0], d["max_u"]), max(worst[1], drift))
        assert d["all_finite"] and d["time_step"] == 50000, i
        assert d["max_u"] < 0.10, (i, d["max_u"])
        assert drift < 1e-4, (i, drift)
    print(f"\n[batch64 fp32 50k] max |u| {worst[0]:.4f}, max mass drift {worst[1]:.2e}")


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16_offset"])
def test_batch_equals_single_cavities(storage):
    batch = CavityBatch(SPECS, storage=storage)
    singles = [single(sp, storage) for sp in SPECS]
    for chunk in (1, 150, 300, 7, 3):   # spin-up direct launches, graph replays, remainders
        batch.run(chunk)
        for s in singles:
            s.run(chunk)
        for i, (c, s) in enumerate(zip(batch, singles)):
            assert c.time_step == s.time_step
            bad = same_state(c, s)
            assert bad is None, f"cavity {i} differs in {bad} after {c.time_step} steps"


def test_autotune_keeps_state_and_picks_a_candidate():
    batch = CavityBatch(SPECS, storage="fp32")
    ref = [single(sp, "fp32") for sp in SPECS]
    batch.run(250)
    for s in ref:
        s.run(250)
    res = batch.autotune(trial_steps=130)
    assert res["candidates"] == 8 + 2 * 3   # 4 shapes x 2 orders, N = 2, 3, 4 x 2 orders
    if res["steps_per_launch"] == 1:
        assert (res["tile_x"], res["tile_y"]) in {(64, 4), (64, 2), (64, 8), (128, 4)}
    else:
        assert (res["steps_per_launch"], res["tile_x"], res["tile_y"]) in {
            (2, 64, 8), (3, 32, 16), (4, 16, 16)}
    assert res["best_ms_per_step"] <= res["first_ms_per_step"]
    for i, (c, s) in enumerate(zip(batch, ref)):
        assert c.time_step == s.time_step
        assert same_state(c, s) is None, f"autotune changed cavity {i}"
    batch.run(200)
    for s in ref:
        s.run(200)
    for i, (c, s) in enumerate(zip(batch, ref)):
        assert same_state(c, s) is None, f"cavity {i} after tuned run"
    d = batch[1].diagnostics()
    assert d["all_finite"] and d["time_step"] == 450


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_batch_with_a_height_outside_the_tiling(storage):
    """ny = 49 fits no N-step tiling (48 = 0 mod 8, 12, 16): the whole batch
    steps one step per launch and still equals the single cavities."""
    specs = SPECS[:2] + [dict(nx=20, ny=49, ra=1e5, cs=0.1)]
    batch = CavityBatch(specs, storage=storage)
    singles = [single(sp, storage) for sp in specs]
    for chunk in (1, 130, 5):
        batch.run(chunk)
        for s in singles:
            s.run(chunk)
        for i, (c, s) in enumerate(zip(batch, singles)):
            bad = same_state(c, s)
            assert bad is None, f"cavity {i} differs in {bad} after {c.time_step} steps"
