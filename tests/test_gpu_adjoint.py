"""Hand-written adjoint of the fused FP64 step: the window objective J against
the oracle (the mean of post.rs's nusselt_hot over the window's steps, with
the same hot-wall profile and Cs), gradients of J with respect to the hot-wall
profile and Cs checked against central finite differences through the
torch.autograd.Function wrapper and against the FP64 tangent, and invariance
under the checkpoint interval."""
import os

import numpy as np
import pytest

from cfd_cavity_amd import CavitySolver
from cfd_cavity_amd.adjoint import NuWindow, adjoint_nu_window, nu_window, tangent_nu_window
from tests.helpers import bits_equal

pytestmark = pytest.mark.gpu


def spun_up(nx, ny, ra, steps, seed=42):
    s = CavitySolver(nx, ny, ra, 0.71, 0.10, 0.9, 0.02, seed, 1)
    s.run(steps)
    return s


def directions(ny, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        dth = rng.standard_normal(ny)
        dth[0] = dth[-1] = 0.0
        out.append((dth / np.linalg.norm(dth), float(rng.standard_normal())))
    return out


def fd_check(s, theta, cs, K, C, dirs, h=1e-6, tol=1e-4):
    import torch

    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    c = torch.tensor(cs, dtype=torch.float64, requires_grad=True)
    j_t = NuWindow().apply(th, c, s, K, C)
    j_t.backward()
    j = float(j_t.detach())
    gth, gcs = th.grad.numpy(), float(c.grad)
    assert gth[0] == 0.0 and gth[-1] == 0.0
    errs = []
    for dth, dcs in dirs:
        jp = nu_window(s, theta + h * dth, cs + h * dcs, K)
        jm = nu_window(s, theta - h * dth, cs - h * dcs, K)
        fd = (jp - jm) / (2 * h)
        ad = float(gth @ dth + gcs * dcs)
        errs.append(abs(ad - fd) / max(abs(fd), 1e-12))
    # the FP64 tangent (forward mode, same per-node templates on dual numbers)
    terrs = []
    for dth, dcs in dirs:
        jt, djt = tangent_nu_window(s, theta, cs, dth, dcs, K)
        assert jt == float(j)  # same forward arithmetic, bitwise
        ad = float(gth @ dth + gcs * dcs)
        terrs.append(abs(ad - djt) / max(abs(djt), 1e-12))
    print(f"\n[adjoint] J={float(j):.10f} |dJ/dtheta|={np.linalg.norm(gth):.4e} dJ/dCs={gcs:.4e} "
          f"max rel err over {len(dirs)} directions: vs FD {max(errs):.2e}, vs tangent "
          f"{max(terrs):.2e}")
    assert max(errs) <= tol, errs
    assert max(terrs) <= 1e-8, terrs
    return float(j), gth, gcs


WINDOW_CASES = [
    # nx, ny, Ra, steps before the window, K, checkpoint interval, Cs, profile
    (16, 40, 1e5, 200, 61, 16, 0.10, False),
    (24, 64, 1e6, 230, 45, 7, 0.13, True),
    (17, 35, 2e5, 90, 38, 5, 0.08, True),       # window starts inside the spin-up
    (96, 384, 1e7, 8000, 2000, 100, 0.10, True),  # the adjoint configuration
]


@pytest.mark.parametrize("nx,ny,ra,t0,K,C,cs,profile", WINDOW_CASES)
def test_window_objective_matches_oracle(oracle, nx, ny, ra, t0, K, C, cs, profile):
    """J = (1/K) sum_{m=1..K} nusselt_hot(S_m) (post.rs:68-78): the oracle
    steps K times from the same state S_0 with the hot-wall profile and Cs set
    and averages its nusselt_hot after each step.  The sums run in another
    order (rows per step there, steps per row here), hence 1e-12 relative.
    The window's J, the adjoint's and the tangent's are the same number.  The
    mean over S_0..S_{K-1} (a window shifted by one step) must miss by more
    than the tolerance, so an off-by-one in the window cannot pass."""
    s = spun_up(nx, ny, ra, t0)
    jobs = min(16, os.cpu_count() or 1)
    o = oracle.OracleSolver(nx, ny, ra, 0.71, 0.10, 0.9, 0.02, 42, jobs)
    o.run(t0)
    theta = (0.5 + 0.01 * np.random.default_rng(nx * ny).uniform(-1, 1, ny)) if profile else None
    o.set_hot_wall(theta)
    o.set_cs(cs)
    nus = [o.diagnostics()["nusselt_hot"]]
    for _ in range(K):
        o.step()
        nus.append(o.diagnostics()["nusselt_hot"])
    j_oracle = sum(nus[1:]) / K
    j_shifted = sum(nus[:-1]) / K
    j = nu_window(s, theta, cs, K)
    print(f"\n[window] {nx}x{ny} t0={t0} K={K}: J={j:.15g} oracle={j_oracle:.15g} "
          f"rel={abs(j - j_oracle) / abs(j_oracle):.2e} shifted rel="
          f"{abs(j - j_shifted) / abs(j_oracle):.2e}")
    assert abs(j - j_oracle) <= 1e-12 * abs(j_oracle)
    assert abs(j - j_shifted) > 1e-12 * abs(j_oracle)
    assert adjoint_nu_window(s, theta, cs, K, C)[0] == j
    assert tangent_nu_window(s, theta, cs, None, 0.0, K)[0] == j
    assert s.time_step == t0


def test_adjoint_small_vs_finite_differences():
    s = spun_up(16, 40, 1e5, 400)
    ny = s.ny()
    rng = np.random.default_rng(5)
    theta = 0.5 + 0.01 * rng.uniform(-1, 1, ny)
    t_before = s.get_field("temperature")
    fd_check(s, theta, 0.10, 60, 16, directions(ny, 8, 9))
    # the solver's own state is untouched by the window evaluations
    assert bits_equal(s.get_field("temperature"), t_before) and s.time_step == 400


def test_adjoint_checkpoint_interval_invariance():
    s = spun_up(24, 64, 1e6, 300)
    theta = np.full(s.ny(), 0.5)
    a = adjoint_nu_window(s, theta, 0.1, 45, 7)
    b = adjoint_nu_window(s, theta, 0.1, 45, 100)
    assert a[0] == b[0] and a[2] == b[2] and bits_equal(a[1], b[1])
    assert a[0] == nu_window(s, theta, 0.1, 45)


def test_adjoint_96x384_ra1e7_window_8k_10k():
    """96x384, Ra=1e7, FP64: J = Nu_hot averaged over steps 8k-10k from the
    seeded spun-up state; theta = 0.5 + seeded +-0.01; Cs = 0.10; checkpoint
    every 100 steps; central differences h = 1e-6 on 16 seeded directions."""
    s = spun_up(96, 384, 1e7, 8000)
    rng = np.random.default_rng(42)
    theta = 0.5 + 0.01 * rng.uniform(-1, 1, s.ny())
    fd_check(s, theta, 0.10, 2000, 100, directions(s.ny(), 16, 1234))


def test_window_buffers_reused_across_calls():
    """The window calls keep their device buffers in the solver between calls:
    repeated calls, calls of other shapes (window length, checkpoint interval)
    and of the other entry points in between give bitwise the results of a
    fresh solver, and the accumulators start from zero on every call."""
    theta = 0.5 + 0.01 * np.random.default_rng(3).uniform(-1, 1, 40)
    dth, dcs = directions(40, 1, 8)[0]
    fresh = {}
    for key, f in [("adj", lambda s: adjoint_nu_window(s, theta, 0.1, 50, 9)),
                   ("adj_small", lambda s: adjoint_nu_window(s, theta, 0.1, 21, 4)),
                   ("tan", lambda s: tangent_nu_window(s, theta, 0.1, dth, dcs, 50)),
                   ("tan_zero", lambda s: tangent_nu_window(s, theta, 0.1, None, 0.0, 33)),
                   ("j", lambda s: nu_window(s, theta, 0.1, 64))]:
        fresh[key] = f(spun_up(16, 40, 1e5, 200))
    s = spun_up(16, 40, 1e5, 200)
    order = ["adj", "adj", "j", "adj_small", "tan", "adj", "tan_zero", "j", "adj_small", "tan"]
    calls = {"adj": lambda: adjoint_nu_window(s, theta, 0.1, 50, 9),
             "adj_small": lambda: adjoint_nu_window(s, theta, 0.1, 21, 4),
             "tan": lambda: tangent_nu_window(s, theta, 0.1, dth, dcs, 50),
             "tan_zero": lambda: tangent_nu_window(s, theta, 0.1, None, 0.0, 33),
             "j": lambda: nu_window(s, theta, 0.1, 64)}
    for key in order:
        got, want = calls[key](), fresh[key]
        if key.startswith("adj"):
            assert got[0] == want[0] and got[2] == want[2] and bits_equal(got[1], want[1]), key
        else:
            assert got == want, key
    assert fresh["tan_zero"][1] == 0.0


@pytest.mark.parametrize("nx,ny", [(3, 40), (17, 35), (40, 33), (20, 49)])
def test_window_blocked_steps_equal_single_steps(nx, ny):
    """The forward window advances up to four steps per launch, each adding the
    Nusselt rows of the state it pulls (blocks end on checkpoint steps); J and
    the adjoint's gradients equal those of one step per launch bitwise.  The
    heights cover every fit of the tilings: 40 and 35 take four steps, 33 two
    and 49 none."""
    s = spun_up(nx, ny, 1e5, 150)
    theta = 0.5 + 0.01 * np.random.default_rng(7).uniform(-1, 1, ny)
    blocked = (nu_window(s, theta, 0.1, 37), adjoint_nu_window(s, theta, 0.1, 37, 6))
    s.time_blocking = 1
    single = (nu_window(s, theta, 0.1, 37), adjoint_nu_window(s, theta, 0.1, 37, 6))
    assert blocked[0] == single[0] == blocked[1][0] == single[1][0]
    assert blocked[1][2] == single[1][2] and bits_equal(blocked[1][1], single[1][1])
