"""Two steps per launch (temporal blocking) against one step per launch.

The two-step kernel runs each node's arithmetic of the single-step kernel and
pulls, in its second step, the values the first step would have stored, so
the two run loops must agree bit for bit on populations, fields, fy_node and
nu_t, for FP64 and FP32 storage, through the spin-up (per-step scalars), the
graph replays after it, ragged tile edges, both g-equilibrium orders and a hot
wall profile.  The FP64 run loop with two steps per launch is also checked
against the oracle by the parity tests (its default).
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


# (nx, ny): the tiles (32x12 in FP64, 64x8 in FP32) with ragged right and top
# edges, the top wall in the second row of its tile (ny - 1 = 1 mod 12 or
# mod 8), a lattice narrower than one tile.
SIZES = [(16, 40), (33, 50), (40, 14), (35, 34), (3, 3), (70, 82)]


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
@pytest.mark.parametrize("nx,ny", SIZES)
def test_two_step_launches_equal_single_steps(storage, nx, ny):
    cfg = dict(SMALL, nx=nx, ny=ny)
    a = make(cfg, 2, storage=storage)
    b = make(cfg, 1, storage=storage)
    assert a.steps_per_launch == 2 and b.steps_per_launch == 1
    print(f"[time blocking] {storage} {nx}x{ny}", flush=True)
    for chunk in (1, 2, 7, 40):
        a.run(chunk)
        b.run(chunk)
        assert_same(a, b, f"{storage} {nx}x{ny} t={a.time_step}")
    a.close()
    b.close()


@pytest.mark.parametrize("kw", [dict(geq_order=1), dict(nu_start_mult=4.0),
                                dict(storage="fp32", nu_start_mult=4.0)])
def test_two_step_graphs_and_spin_up(kw):
    """Long runs: direct pairs through the spin-up, then graph replays of
    32 two-step launches, then the direct tail that writes nu_t."""
    a = make(SMALL, 2, **kw)
    b = make(SMALL, 1, **kw)
    print(f"[time blocking] graphs {kw}", flush=True)
    for chunk in (301, 130, 3):
        a.run(chunk)
        b.run(chunk)
        assert_same(a, b, f"{kw} t={a.time_step}")


def test_two_step_hot_wall_profile():
    rng = np.random.default_rng(5)
    theta = 0.5 + 0.01 * rng.uniform(-1, 1, SMALL["ny"])
    a = make(SMALL, 2)
    b = make(SMALL, 1)
    a.set_hot_wall(theta)
    b.set_hot_wall(theta)
    a.run(157)
    b.run(157)
    assert_same(a, b, "hot wall profile")


def test_two_step_switching_mid_run():
    """Changing the blocking between runs keeps the state and the fy planes."""
    a = make(SMALL, 2)
    b = make(SMALL, 1)
    for k, chunk in enumerate((5, 9, 130, 4, 77)):
        a.time_blocking = 1 + (k % 2)
        a.run(chunk)
        b.run(chunk)
        assert_same(a, b, f"switch t={a.time_step}")


@pytest.mark.parametrize("storage,ny", [("fp64", 37), ("fp32", 33)])
def test_heights_outside_the_tiling_use_single_steps(storage, ny):
    cfg = dict(SMALL, nx=20, ny=ny)  # (ny - 1) % 12 == 0 / (ny - 1) % 8 == 0
    a = make(cfg, 2, storage=storage)
    assert a.steps_per_launch == 1
    b = make(cfg, 1, storage=storage)
    a.run(50)
    b.run(50)
    assert_same(a, b, f"{storage} 20x{ny}")


def test_fp16_offset_keeps_single_steps():
    a = make(SMALL, 2, storage="fp16_offset")
    assert a.steps_per_launch == 1
