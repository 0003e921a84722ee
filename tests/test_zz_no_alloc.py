"""The allocation-free window objective (cavlbm_nu_window_static): through the
C-ABI, J equals cavlbm_nu_window's bit for bit on two calls in a row, the
second on a smaller lattice and a shorter window (the static array keeps the
first call's contents), then on the first again; the solver's state is left
unchanged; a lattice beyond the array's limit is a size error."""
import numpy as np
import pytest

from cfd_cavity_amd import CavitySolver
from cfd_cavity_amd._lib import E_SIZE, CavlbmError
from cfd_cavity_amd.adjoint import nu_window, nu_window_static
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu


def spun_up(nx, ny, ra, steps):
    s = CavitySolver(nx, ny, ra, 0.71, 0.10, 0.9, 0.02, 42, 1)
    s.run(steps)
    return s


def test_static_window_equals_allocating_window():
    a = spun_up(96, 384, 1e7, 400)
    theta = 0.5 + 0.01 * np.random.default_rng(11).uniform(-1, 1, a.ny())
    t_before = a.get_field("temperature")
    j_static = nu_window_static(a, theta, 0.10, 300)
    j_alloc = nu_window(a, theta, 0.10, 300)
    print(f"\n[no-alloc] 96x384 K=300: static {j_static!r} allocating {j_alloc!r}")
    assert j_static == j_alloc
    b = spun_up(24, 64, 1e6, 230)
    j2_static = nu_window_static(b, None, 0.13, 45)
    j2_alloc = nu_window(b, None, 0.13, 45)
    print(f"[no-alloc] 24x64 K=45: static {j2_static!r} allocating {j2_alloc!r}")
    assert j2_static == j2_alloc
    assert nu_window_static(a, theta, 0.10, 300) == j_alloc
    assert a.time_step == 400 and b.time_step == 230
    assert bits_equal(a.get_field("temperature"), t_before)


def test_static_window_rejects_a_lattice_beyond_the_array():
    s = CavitySolver(256, 1024, 1e10, 0.71, 0.10, 0.9, 0.005, 42, 1)   # plane 262144 nodes
    with pytest.raises(CavlbmError) as e:
        nu_window_static(s, None, 0.10, 10)
    assert e.value.code == E_SIZE


def test_static_window_at_the_array_limits():
    """64x1024 fills the static array exactly (a pitched plane of 65 536 nodes
    and 1024 rows): J equals the allocating entry point's.  One column or one
    row more is a size error."""
    s = spun_up(64, 1024, 1e6, 60)
    theta = 0.5 + 0.01 * np.random.default_rng(5).uniform(-1, 1, 1024)
    j_static = nu_window_static(s, theta, 0.10, 24)
    j_alloc = nu_window(s, theta, 0.10, 24)
    print(f"\n[no-alloc] 64x1024 K=24: static {j_static!r} allocating {j_alloc!r}")
    assert j_static == j_alloc
    for nx, ny in ((65, 1024), (64, 1025)):
        t = CavitySolver(nx, ny, 1e6, 0.71, 0.10, 0.9, 0.02, 42, 1)
        with pytest.raises(CavlbmError) as e:
            nu_window_static(t, None, 0.10, 10)
        assert e.value.code == E_SIZE, (nx, ny)
