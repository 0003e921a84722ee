"""Several steps per launch (temporal blocking) against one step per launch.

The N-step kernels run each node's arithmetic of the single-step kernel and
pull, in each later sub-step, the values the previous step would have
stored, so the run loops must agree bit for bit on populations, fields,
fy_node and nu_t, for FP64 and FP32 storage and N = 2, 3, 4, through the
spin-up (per-step scalars), the graph replays after it, ragged tile edges,
both g-equilibrium orders and a hot wall profile.  The FP64 run loop with its
default four steps per launch is also checked against the oracle by the
parity tests.
"""
import numpy as np
import pytest

from cfd_cavity_amd import CavitySolver
from tests.helpers import SMALL, args_of, bits_equal, first_mismatch

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "ux", "uy", "temperature", "nu_t", "fy_node")


def make(cfg, tblock, **kw):
    s = CavitySolver(*args_of(cfg), **kw)
    s.time_blocking = tblock
    return s


def assert_same(a, b, what):
    for name in FIELDS:
        x, y = a.get_field(name), b.get_field(name)
        assert bits_equal(x, y), f"{what} {name}: {first_mismatch(x, y)}"
    for which in ("f", "g"):
        x, y = a.get_pops(which), b.get_pops(which)
        assert bits_equal(x, y), f"{what} pops {which}: {first_mismatch(x, y)}"


# Output tiles: FP64 32x12 (N=2), 16x16 (N=3, 4); FP32 64x8 (N=2), 32x16
# (N=3), 16x16 (N=4).  (nx, ny): ragged right and top tile edges; ny = 50
# puts the top wall in the second row of its tile for all of them
# (49 = 1 mod 8, 12, 16), as ny = 14 does for 32x12 and ny = 34 for 16-row
# tiles; a lattice narrower than one tile.
SIZES = [(16, 40), (33, 50), (40, 14), (35, 34), (3, 3), (70, 82)]


def fits(storage, n, ny):
    ty = {("fp64", 2): 12, ("fp64", 3): 16, ("fp64", 4): 16,
          ("fp32", 2): 8, ("fp32", 3): 16, ("fp32", 4): 16}[
        ("fp32" if storage == "fp16_offset" else storage, n)]
    return (ny - 1) % ty != 0


@pytest.mark.parametrize("n", [2, 3, 4])
@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16_offset"])
@pytest.mark.parametrize("nx,ny", SIZES)
def test_block_launches_equal_single_steps(storage, n, nx, ny):
    cfg = dict(SMALL, nx=nx, ny=ny)
    a = make(cfg, n, storage=storage)
    b = make(cfg, 1, storage=storage)
    expect = max([k for k in range(2, n + 1) if fits(storage, k, ny)], default=1)
    assert a.steps_per_launch == expect and b.steps_per_launch == 1
    print(f"[time blocking] {storage} n={n} {nx}x{ny}", flush=True)
    for chunk in (1, 2, 3, 5, 7, 40):
        a.run(chunk)
        b.run(chunk)
        assert_same(a, b, f"{storage} n={n} {nx}x{ny} t={a.time_step}")
    a.close()
    b.close()


def test_defaults():
    assert make(SMALL, 4).steps_per_launch == 4
    a = CavitySolver(*args_of(SMALL))
    assert a.steps_per_launch == 4  # FP64 default
    a = CavitySolver(*args_of(SMALL), storage="fp32")
    assert a.steps_per_launch == 3
    a = CavitySolver(*args_of(SMALL), storage="fp16_offset")
    assert a.steps_per_launch == 3
    a = CavitySolver(*args_of(dict(SMALL, ny=17)))  # 16 % 16 == 0: N=3, 4 do not fit
    assert a.steps_per_launch == 2
    with pytest.raises(Exception):
        a.time_blocking = 5


@pytest.mark.parametrize("n,kw", [(4, dict(geq_order=1)), (4, dict(nu_start_mult=4.0)),
                                  (3, dict(nu_start_mult=4.0)), (2, dict(nu_start_mult=4.0)),
                                  (3, dict(storage="fp32", nu_start_mult=4.0)),
                                  (3, dict(storage="fp16_offset", nu_start_mult=4.0))])
def test_block_graphs_and_spin_up(n, kw):
    """Long runs: direct blocks through the spin-up, then graph replays of
    48 steps, then the direct tail that writes nu_t."""
    a = make(SMALL, n, **kw)
    b = make(SMALL, 1, **kw)
    print(f"[time blocking] graphs n={n} {kw}", flush=True)
    for chunk in (301, 130, 3):
        a.run(chunk)
        b.run(chunk)
        assert_same(a, b, f"{kw} t={a.time_step}")


def test_block_hot_wall_profile():
    rng = np.random.default_rng(5)
    theta = 0.5 + 0.01 * rng.uniform(-1, 1, SMALL["ny"])
    a = make(SMALL, 4)
    b = make(SMALL, 1)
    a.set_hot_wall(theta)
    b.set_hot_wall(theta)
    a.run(157)
    b.run(157)
    assert_same(a, b, "hot wall profile")


def test_block_switching_mid_run():
    """Changing the blocking between runs keeps the state and the fy planes."""
    a = make(SMALL, 2)
    b = make(SMALL, 1)
    for k, chunk in enumerate((5, 9, 130, 4, 77, 101, 6)):
        a.time_blocking = (2, 4, 1, 3)[k % 4]
        a.run(chunk)
        b.run(chunk)
        assert_same(a, b, f"switch t={a.time_step}")


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16_offset"])
def test_heights_outside_the_tiling_use_single_steps(storage):
    ny = 49  # 48 = 0 mod 8, 12 and 16: no block fits
    cfg = dict(SMALL, nx=20, ny=ny)
    a = make(cfg, 4, storage=storage)
    assert a.steps_per_launch == 1
    b = make(cfg, 1, storage=storage)
    a.run(50)
    b.run(50)
    assert_same(a, b, f"{storage} 20x{ny}")


def test_fp16_offset_hot_wall_blocks():
    """The FP16-offset g deviations refer to the writing node's temperature;
    with a hot-wall profile the rounding between sub-steps must still follow
    the stored temperature plane exactly."""
    rng = np.random.default_rng(9)
    theta = 0.5 + 0.01 * rng.uniform(-1, 1, SMALL["ny"])
    a = make(SMALL, 3, storage="fp16_offset")
    b = make(SMALL, 1, storage="fp16_offset")
    a.set_hot_wall(theta)
    b.set_hot_wall(theta)
    a.run(131)
    b.run(131)
    assert_same(a, b, "fp16 hot wall")
