"""Pins of the CPU oracle (no GPU): the reference's own tests restated
(src/uniform/tests.rs), closed forms, invariants, and an independent numpy
textbook step written in a different algebraic form (tests/textbook_d2q9.py).

Each check is chosen so that a plausible mistake in the oracle (a dropped term,
a wrong sign, a wrong opposite/direction index, a transposed velocity
component, the wrong relaxation time on a force part, the wrong BC order)
fails at least one of them.
"""
import math

import numpy as np
import pytest

from tests.helpers import SMALL, args_of, bits_equal
from tests.textbook_d2q9 import TextbookCavity

CS2 = 1.0 / 3.0


def small(oracle, **kw):
    return oracle.OracleSolver(*args_of(SMALL), 1, **kw)


# ---------------------------------------------------------------------------
# lattice and equilibria: closed forms

def test_lattice_tables(oracle):
    cx, cy, opp, w = oracle.lattice()
    c = np.stack([cx, cy], 1)
    assert abs(w.sum() - 1.0) < 1e-12                               # tests.rs:130-133
    for i in range(9):
        assert (c[opp[i]] == -c[i]).all() and opp[opp[i]] == i
    # isotropy: sum w c_a c_b = cs2 delta_ab, 4th moment sum w cx^2 cy^2 = cs2^2
    assert abs((w * cx * cx).sum() - CS2) < 1e-15 and abs((w * cx * cy).sum()) < 1e-15
    assert abs((w * cx * cx * cy * cy).sum() - CS2 * CS2) < 1e-15
    assert sorted(opp[1:5]) == [5, 6, 7, 8]                         # swap_opposite i = 1..4


@pytest.mark.parametrize("rho,ux,uy", [(1.3, 0.07, -0.03), (0.97, -0.02, 0.05), (1.0, 0.0, 0.0)])
def test_feq_moments_closed_form(oracle, rho, ux, uy):
    cx, cy, _, _ = oracle.lattice()
    f = np.array([oracle.feq(i, rho, ux, uy) for i in range(9)])
    assert abs(f.sum() - rho) < 1e-12                                # tests.rs:116-120
    assert abs((cx * f).sum() - rho * ux) < 1e-14
    assert abs((cy * f).sum() - rho * uy) < 1e-14
    assert abs((cx * cx * f).sum() - rho * (CS2 + ux * ux)) < 1e-14
    assert abs((cx * cy * f).sum() - rho * ux * uy) < 1e-14
    assert abs((cy * cy * f).sum() - rho * (CS2 + uy * uy)) < 1e-14


@pytest.mark.parametrize("order", [1, 2])
def test_geq_moments_closed_form(oracle, order):
    cx, cy, _, _ = oracle.lattice()
    t, ux, uy = 0.37, 0.05, 0.02
    g = np.array([oracle.geq(i, t, ux, uy, order) for i in range(9)])
    assert abs(g.sum() - t) < 1e-12                                  # tests.rs:123-127
    assert abs((cx * g).sum() - t * ux) < 1e-14
    assert abs((cy * g).sum() - t * uy) < 1e-14
    for i in range(9):  # wall value geq(T, 0, 0) = w_i T exactly
        assert oracle.geq(i, -0.5, 0.0, 0.0, order) == oracle.lattice()[3][i] * -0.5


# ---------------------------------------------------------------------------
# collide_node: closed forms

def test_collide_from_equilibrium_force_moments(oracle):
    """From f = feq: mass kept, x-momentum kept, y-momentum gains
    (1 - 1/(2 tau-)) * fy (the odd Guo part), g unchanged, nu_t = 0."""
    cx, cy, _, _ = oracle.lattice()
    rho, ux, uy, t = 1.02, 0.013, -0.021, 0.31
    tau, beta, nu, alpha = 0.53, 3e-4, 0.01, 0.014
    f0 = np.array([oracle.feq(i, rho, ux, uy) for i in range(9)])
    g0 = np.array([oracle.geq(i, t, ux, uy) for i in range(9)])
    f, g, nut, fy = oracle.collide_node(f0, g0, rho, ux, uy, t, tau, beta, 0.0, 0.1, 0.9, nu, alpha)
    assert fy == rho * beta * t
    assert nut == 0.0
    tp = 0.5 + 3.0 * nu
    tm = 0.5 + (3.0 / 16.0) / (tp - 0.5)
    assert abs(f.sum() - rho) < 1e-15
    assert abs((cx * f).sum() - rho * ux) < 1e-15
    assert abs((cy * f).sum() - (rho * uy + (1.0 - 0.5 / tm) * fy)) < 1e-15
    assert np.allclose(g, g0, rtol=0, atol=1e-16)


def test_collide_smagorinsky_closed_form(oracle):
    """A pure shear non-equilibrium eps*w_i*cx*cy/(cs2^2) gives pi_xy = eps,
    |S| = sqrt(2) eps / (2 cs2 tau rho), nu_t = Cs^2 |S|, and relaxes with the
    even rate 1/tau+ (the shear moment is even)."""
    cx, cy, _, w = oracle.lattice()
    rho, eps, tau, cs = 1.0, 1e-4, 0.52, 0.1
    feq = np.array([oracle.feq(i, rho, 0.0, 0.0) for i in range(9)])
    f0 = feq + eps * w * cx * cy / (CS2 * CS2)
    pxy = (cx * cy * (f0 - feq)).sum()
    assert abs(pxy - eps) < 1e-12 * eps
    g0 = np.zeros(9)
    f, _, nut, _ = oracle.collide_node(f0, g0, rho, 0.0, 0.0, 0.0, tau, 0.0, 0.0, cs, 0.9, 0.01, 0.01)
    smag = math.sqrt(2.0) * pxy / (2.0 * CS2 * tau * rho)
    assert abs(nut - cs * cs * smag) < 1e-12 * nut
    tp = 0.5 + 3.0 * (0.01 + nut)
    assert abs((cx * cy * (f - feq)).sum() - pxy * (1.0 - 1.0 / tp)) < 1e-12 * eps


def test_poiseuille_closed_form(oracle):
    """tests.rs:246-376 with the force along y: the transient first-mode
    Poiseuille solution; a force factor on the wrong relaxation time gives
    ~0.05x the target."""
    nx, ny, nu, fy, n = 21, 30, 0.01, 1.0e-6, 15_000
    h = float(nx)
    u_steady = fy * h * h / (8.0 * nu)
    conv = 1.0 - math.exp(-n * nu * math.pi ** 2 / (h * h))
    u = oracle.poiseuille_channel(nx, ny, nu, fy, n)
    assert abs(u - u_steady * conv) / u_steady < 0.02


# ---------------------------------------------------------------------------
# calibration (solver.rs:123-154, :194) against the closed-form table

@pytest.mark.parametrize("nx,ny,ra,ut,nu,alpha,beta,taup,spin", [
    (64, 256, 1e6, 0.02, 0.018557, 0.026136, 2.925e-5, 0.55567, 200),
    (96, 384, 1e7, 0.02, 0.011296, 0.015910, 3.199e-5, 0.53389, 200),
    (128, 512, 1e8, 0.02, 0.006355, 0.008950, 4.262e-5, 0.51906, 200),
    (256, 1024, 1e10, 0.02, 0.002175, 0.003063, 6.222e-5, 0.50652, 200),
    (256, 1024, 1e10, 0.005, 0.002, 0.002817, 5.262e-5, 0.506, 200),
])
def test_calibration_table(oracle, nx, ny, ra, ut, nu, alpha, beta, taup, spin):
    o = oracle.OracleSolver(nx, ny, ra, 0.71, 0.1, 0.9, ut, 42, 1)
    d = o.derived()
    lz = ny - 1
    nu_cf = max((ut * math.sqrt(0.71) * lz / (1.7 * math.sqrt(ra))) ** (2.0 / 3.0), 0.002)
    assert abs(d["nu"] - nu_cf) <= 1e-15 * nu_cf
    assert abs(d["nu"] - nu) / nu < 2e-4
    assert abs(d["alpha"] - alpha) / alpha < 2e-4
    assert abs(d["beta_eff"] - beta) / beta < 2e-4
    assert abs(d["tau"] - taup) < 1e-5
    assert d["spin_up_steps"] == spin
    assert d["beta_eff"] == ra * d["nu"] * d["alpha_phys"] / lz ** 3


# ---------------------------------------------------------------------------
# init and RNG (solver.rs:160-176, rng.rs)

def test_init_state(oracle):
    o = small(oracle)
    nx, ny = SMALL["nx"], SMALL["ny"]
    t = o.temperature.reshape(ny, nx)
    lin = 0.5 - np.arange(nx) / (nx - 1)
    pert = t - lin[None, :]
    assert np.all(np.abs(pert) <= 1e-3)                              # tests.rs:138-149
    assert np.all(o.rho == 1.0) and np.all(o.ux == 0.0) and np.all(o.uy == 0.0)
    _, _, _, w = oracle.lattice()
    assert np.all(o.f == w[None, :])
    assert np.allclose(o.g, w[None, :] * o.temperature[:, None], rtol=0, atol=1e-17)


def test_rng_xorshift64star_properties(oracle):
    import ctypes as C

    L = oracle.lib()
    st = C.c_uint64(42 | 1)
    xs = np.array([L.orc_rng_next_f64(C.byref(st)) for _ in range(20000)])
    assert xs.min() >= 0.0 and xs.max() < 1.0
    assert len(np.unique(xs)) == xs.size
    # uniformity (Kolmogorov-Smirnov against U[0,1))
    from scipy import stats

    assert stats.kstest(xs, "uniform").pvalue > 1e-3
    # the init draws the noise in (iy outer, ix inner) order from seed|1
    o = small(oracle)
    st = C.c_uint64(42 | 1)
    u = np.array([L.orc_rng_next_f64(C.byref(st)) for _ in range(SMALL["nx"] * SMALL["ny"])])
    nx = SMALL["nx"]
    lin = 0.5 + (-1.0) * ((np.arange(nx * SMALL["ny"]) % nx) / (nx - 1))
    assert bits_equal(o.temperature, lin + 1.0e-3 * (2.0 * u - 1.0))


# ---------------------------------------------------------------------------
# independent textbook step

def textbook_of(o, **kw):
    nx, ny = o.nx, o.ny
    d = o.derived()

    def planes(a):
        return a.reshape(ny, nx, 9).transpose(2, 0, 1).copy()

    def fld(a):
        return a.reshape(ny, nx).copy()

    return TextbookCavity(planes(o.f), planes(o.g), fld(o.rho), fld(o.ux), fld(o.uy),
                          fld(o.temperature), fld(o.fy_node), fld(o.nu_t), nu=d["nu"],
                          alpha=d["alpha"], beta_eff=d["beta_eff"], spin_up=d["spin_up_steps"],
                          cs=d["cs"], pr_t=d["pr_t"], t_ref=d["t_ref"], **kw)


def compare_textbook(o, tb, tol):
    nx, ny = o.nx, o.ny
    f = o.f.reshape(ny, nx, 9).transpose(2, 0, 1)
    g = o.g.reshape(ny, nx, 9).transpose(2, 0, 1)
    assert np.max(np.abs(f - tb.f)) < tol
    assert np.max(np.abs(g - tb.g)) < tol
    for name, ref in (("rho", tb.rho), ("ux", tb.ux), ("uy", tb.uy), ("temperature", tb.temp),
                      ("fy_node", tb.fy), ("nu_t", tb.nu_t)):
        got = getattr(o, name).reshape(ny, nx)
        assert np.max(np.abs(got - ref)) < tol, name


@pytest.mark.parametrize("cfg,nsteps", [(SMALL, 1), (SMALL, 30), (SMALL, 230),
                                        (dict(SMALL, nx=12, ny=20, ra=1e6), 60)])
def test_oracle_matches_textbook(oracle, cfg, nsteps):
    o = oracle.OracleSolver(*args_of(cfg), 1)
    tb = textbook_of(o)
    for _ in range(nsteps):
        o.step()
        tb.step()
    compare_textbook(o, tb, 1e-12)


def test_oracle_matches_textbook_variants(oracle):
    o = oracle.OracleSolver(*args_of(SMALL), 1, geq_order=1)
    rng = np.random.default_rng(3)
    theta = 0.5 + 0.01 * rng.uniform(-1, 1, SMALL["ny"])
    o.set_hot_wall(theta)
    tb = textbook_of(o, theta_hot=theta, quadratic_g=False)
    for _ in range(40):
        o.step()
        tb.step()
    compare_textbook(o, tb, 1e-12)


def test_textbook_detects_planted_mistake(oracle):
    """The textbook comparison is sensitive: one step with a plausible mistake
    (force sign flipped, or the odd Guo part relaxed with tau+ instead of
    tau-) departs from the oracle far beyond the 1e-12 tolerance."""
    import tests.textbook_d2q9 as T

    o = oracle.OracleSolver(*args_of(SMALL), 1)
    o.run(210)
    f_before = o.f.copy()
    o.step()
    nx, ny = o.nx, o.ny
    f_after = o.f.reshape(ny, nx, 9).transpose(2, 0, 1)

    def one_step(mutate):
        o2 = oracle.OracleSolver(*args_of(SMALL), 1)
        o2.run(210)
        assert bits_equal(o2.f, f_before)
        tb = textbook_of(o2)
        tb.time_step = 210
        mutate(tb)
        tb.step()
        return np.max(np.abs(tb.f - f_after))

    assert one_step(lambda tb: None) < 1e-12
    assert one_step(lambda tb: setattr(tb, "beta_eff", -tb.beta_eff)) > 1e-9
    saved = T.OPP.copy()
    try:
        T.OPP[1:] = np.roll(saved[1:], 1)   # a wrong opposite table
        assert one_step(lambda tb: None) > 1e-9
    finally:
        T.OPP[:] = saved


# ---------------------------------------------------------------------------
# invariants

def test_jobs_bitwise(oracle):
    a = oracle.OracleSolver(*args_of(dict(SMALL, nx=32, ny=80)), 1)
    b = oracle.OracleSolver(*args_of(dict(SMALL, nx=32, ny=80)), 4)
    a.run(300)
    b.run(300)
    for n in ("rho", "ux", "uy", "temperature", "nu_t", "fy_node", "f", "g"):
        assert bits_equal(getattr(a, n), getattr(b, n)), n


def test_centro_symmetry(oracle):
    """Noise-free linear T: (x, y, T, u) -> (-x, -y, -T, -u) is a symmetry."""
    nx, ny = 16, 40
    o = oracle.OracleSolver(nx, ny, 1e5, 0.71, 0.1, 0.9, 0.02, 42, 1)
    n = nx * ny
    t = np.array([0.5 - (k % nx) / (nx - 1) for k in range(n)])
    o.init_from_fields(np.ones(n), np.zeros(n), np.zeros(n), t)
    o.run(300)
    T = o.temperature.reshape(ny, nx)
    U = o.ux.reshape(ny, nx)
    V = o.uy.reshape(ny, nx)
    R = o.rho.reshape(ny, nx)
    assert np.max(np.abs(T + T[::-1, ::-1])) < 1e-12
    assert np.max(np.abs(U + U[::-1, ::-1])) < 1e-12
    assert np.max(np.abs(V + V[::-1, ::-1])) < 1e-12
    assert np.max(np.abs(R - R[::-1, ::-1])) < 1e-12
    assert np.max(np.abs(V)) > 1e-4  # the flow is not trivially zero


def test_rest_state_without_buoyancy(oracle):
    """Ra = 0 (beta = 0): f stays at rest, u stays ~0, T diffuses between the walls."""
    o = oracle.OracleSolver(16, 40, 0.0, 0.71, 0.1, 0.9, 0.02, 42, 1, nu_override=0.02)
    _, _, _, w = oracle.lattice()
    o.run(100)
    assert np.max(np.abs(o.f - w[None, :])) < 1e-15
    assert np.max(np.abs(o.ux)) < 1e-15 and np.max(np.abs(o.uy)) < 1e-15


def test_mass_conservation_long(oracle):
    o = oracle.OracleSolver(*args_of(dict(SMALL, nx=32, ny=128, ra=1e7)), 4)
    m0 = o.total_f_mass()
    o.run(2000)
    assert abs(o.total_f_mass() - m0) / m0 < 1e-12
    assert o.diagnostics()["all_finite"]


# ---------------------------------------------------------------------------
# the reference's tests (src/uniform/tests.rs) on the oracle

def test_mass_conserved_per_phase(oracle):            # tests.rs:154-169
    s = small(oracle)
    m0 = s.total_f_mass()
    s.collide(100)
    m1 = s.total_f_mass()
    s.stream()
    m2 = s.total_f_mass()
    s.apply_bounce_back()
    m3 = s.total_f_mass()
    s.apply_thermal_boundary_conditions()
    m4 = s.total_f_mass()
    assert all(abs(m - m0) < 1e-8 for m in (m1, m2, m3, m4))


def test_guo_forcing_consistency(oracle):             # tests.rs:174-222
    s = small(oracle)
    m0 = s.total_f_mass()
    s.run(210)
    assert abs(s.total_f_mass() - m0) / abs(m0) < 1e-9
    cy = oracle.lattice()[1].astype(float)
    nx, ny = s.nx, s.ny
    for iy in range(1, ny - 1):
        for ix in range(1, nx - 1):
            k = s.idx(ix, iy)
            mom = 0.0
            for j in range(9):
                mom += cy[j] * s.f[k, j]
            e = mom / s.rho[k]
            e += 0.5 * s.fy_node[k] / s.rho[k]
            assert abs(s.uy[k] - e) < 1e-12


def test_hot_fluid_pushed_upward(oracle):             # tests.rs:381-401
    s = small(oracle)
    s.run(200)
    vals = [s.uy[s.idx(ix, iy)] for iy in range(s.ny // 2, s.ny - 2) for ix in range(1, 4)]
    assert np.mean(vals) > 0.0


def test_walls_after_50_steps(oracle):                # tests.rs:406-476
    s = small(oracle)
    s.run(50)
    nx, ny = s.nx, s.ny
    t = s.temperature.reshape(ny, nx)
    assert np.all(np.abs(t[:, 0] - 0.5) < 1e-9) and np.all(np.abs(t[:, -1] + 0.5) < 1e-9)
    assert np.all(np.abs(t[0] - t[1]) < 1e-9) and np.all(np.abs(t[-1] - t[-2]) < 1e-9)
    for a in (s.ux.reshape(ny, nx), s.uy.reshape(ny, nx)):
        for edge in (a[:, 0], a[:, -1], a[0], a[-1]):
            assert np.all(np.abs(edge) < 1e-12)


def test_nusselt_conduction_limit(oracle):            # tests.rs:481-506
    s = small(oracle)
    nx, ny = s.nx, s.ny
    for iy in range(ny):
        for ix in range(nx):
            s.temperature[s.idx(ix, iy)] = 0.5 + (-0.5 - 0.5) * (ix / (nx - 1))
    d = s.diagnostics()
    assert abs(d["nusselt_hot"] - (ny - 1) / (nx - 1)) < 0.05
    assert abs(d["nusselt_cold"] - (ny - 1) / (nx - 1)) < 0.05


def test_thermal_wall_link_conductance(oracle):       # tests.rs:602-643
    s = oracle.OracleSolver(21, 21, 1.0, 1.0, 0.0, 1.0, 0.02, 42, 1, nu_override=0.00169)
    s.run(20_000)
    t = s.temperature.reshape(21, 21)
    r = slice(1, 20)
    hot = np.mean(np.abs(t[r, 0] - t[r, 1])) / np.mean(np.abs(t[r, 1] - t[r, 2]))
    cold = np.mean(np.abs(t[r, 20] - t[r, 19])) / np.mean(np.abs(t[r, 19] - t[r, 18]))
    assert abs(hot - 0.5) < 0.05 and abs(cold - 0.5) < 0.05


def test_diagnostics_definitions(oracle):
    """post.rs formulas evaluated directly on the oracle's fields."""
    s = oracle.OracleSolver(*args_of(dict(SMALL, nx=24, ny=64, ra=1e6)), 1)
    s.run(400)
    d = s.diagnostics()
    nx, ny = s.nx, s.ny
    T = s.temperature.reshape(ny, nx)
    V = s.uy.reshape(ny, nx)
    U = s.ux.reshape(ny, nx)
    lz = ny - 1
    assert abs(d["nusselt_hot"] - np.abs(T[:, 1] - T[:, 2]).mean() * lz) < 1e-12
    assert abs(d["nusselt_cold"] - np.abs(T[:, -2] - T[:, -3]).mean() * lz) < 1e-12
    assert abs(d["nusselt_hot_wall_link"] - np.abs(T[:, 0] - T[:, 1]).mean() * lz) < 1e-12
    assert abs(d["nusselt_volume"] - (1 + (V[1:-1, 1:-1] * T[1:-1, 1:-1]).mean() * lz)) < 1e-12
    assert d["max_uy"] == np.abs(V).max() and d["max_u"] == max(np.abs(U).max(), np.abs(V).max())
    nut = s.nu_t.reshape(ny, nx)
    nu = s.derived()["nu"]
    tw = (nu + nut[1:-1, 1]) * np.abs(V[1:-1, 1] - V[1:-1, 0])
    assert abs(d["y_plus_hot"] - np.mean(0.5 * np.sqrt(np.maximum(tw, 0)) / nu)) < 1e-12
    assert abs(d["total_mass"] - s.rho.sum()) < 1e-9
