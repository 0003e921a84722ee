"""Pins of the octree oracle's dynamic mixed-scale SGS (no GPU).

The oracle restates src/octree/smag.rs (the uniform model on each level's
arrays, orc_dynamic_model) and its place in step_cavity (solver.rs:647-947).
Pinned by
  - the reference's V5 tests (src/octree/tests.rs:572-786): the model off
    leaves the path bitwise unchanged; a three-level run stays finite with
    coefficients in [0, 0.04]; the interface factors xi are the molecular ones
    at step 0 and with the model off, follow the effective relaxation times
    with it on, and are deterministic;
  - the coarse level's nu_t / alpha_t against the independent numpy model
    (tests/dynamic_sgs_numpy.py) on the level's fields;
  - the floors of each level, which fix the molecular nu / alpha the model
    receives there (nu, 2 nu, 4 nu; the floor binds at a low viscosity);
  - the per-node collision (relaxation times from nu + nu_t) against an
    independent textbook TRT / Guo / BGK form.
"""
import numpy as np
import pytest

import tests.dynamic_sgs_numpy as DN
from tests.helpers import bits_equal
from tests.textbook_d2q9 import CX, CY, OPP, WW, equilibrium

CS2 = 1.0 / 3.0
MARGIN = (0.505 - 0.5) / 3.0


def small(oracle, levels=2, **kw):
    return oracle.OracleOctree(32, 64, 4, 1.0e6, levels=levels, **kw)   # tests.rs:577-579


def molecular_xi(oracle, nu, alpha, odd_exact=False):
    tc, tf, ts = (np.array(oracle_relax(oracle, m * nu, m * alpha)) for m in (1.0, 2.0, 4.0))
    a, b = np.zeros(3), np.zeros(3)
    L = oracle.lib()
    L.orc_oct_rescaling_factors(oracle._dp(tc), oracle._dp(tf), 2, int(odd_exact), oracle._dp(a))
    L.orc_oct_rescaling_factors(oracle._dp(tf), oracle._dp(ts), 2, int(odd_exact), oracle._dp(b))
    return a, b


def oracle_relax(oracle, nu, alpha):
    out = np.zeros(3)
    oracle.lib().orc_oct_relax_times(nu, alpha, oracle._dp(out))
    return out


def test_v5_model_off_is_the_default_path(oracle):             # tests.rs:601-619
    s, ref = small(oracle), small(oracle)
    s.set_dynamic_sgs(False, 0.5, -0.1, 2.0)
    s.run(5)
    ref.run(5)
    for gs, gr in zip(s.grids, ref.grids):
        assert bits_equal(gs.f, gr.f) and bits_equal(gs.g, gr.g)
        assert bits_equal(gs.temperature, gr.temperature)
    assert s.dynamic_stats() == (0.0,) * 5


def test_v5_model_on_runs_sane(oracle):                       # tests.rs:627-673
    sim = small(oracle, levels=3)
    sim.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    for _ in range(5):
        sim.step()
        cs_mean, cs_max, k_mean, k_max, nu_t_max = sim.dynamic_stats()
        assert np.isfinite(cs_mean) and np.isfinite(cs_max)
        assert 0.0 <= cs_max <= 0.04 and cs_mean <= cs_max + 1e-12
        assert 0.0 <= k_max <= 0.04
        for g in sim.grids[:2]:
            for a in (g.rho, g.ux, g.uy, g.temperature):
                assert np.all(np.isfinite(a))
        for gi in (0, 1, 3):
            assert np.all(np.isfinite(sim.sgs_pre(gi, 0)))
            assert np.all(np.isfinite(sim.sgs_pre(gi, 1)))
        assert sim.check_guards() == 0


def test_v5_xi_follows_the_effective_relaxation_times(oracle):  # tests.rs:686-786
    d = small(oracle).derived()
    nu, alpha = d["nu_lu"], d["alpha_lu"]
    xi_mol, xi_sub_mol = molecular_xi(oracle, nu, alpha)
    # 1. step 0: the molecular factors, bitwise
    s0 = small(oracle)
    assert bits_equal(s0.xi()[:3], xi_mol)
    # 2. model on: xi follows nu + mean nu_t, close to the molecular one
    sim = small(oracle, levels=3)
    sim.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    sim.run(5)
    assert sim.dynamic_stats()[4] > 0.0
    xi = sim.xi()
    assert xi[0] != xi_mol[0] and xi[2] != xi_mol[2]
    assert xi[0] > 0.0 and abs(xi[0] - xi_mol[0]) / xi_mol[0] < 0.1
    assert xi[3] != xi_sub_mol[0]                 # interface B of the hot block
    # 3. model off: unchanged after steps
    off = small(oracle, levels=3)
    off.run(5)
    assert bits_equal(off.xi()[:3], xi_mol)
    assert bits_equal(off.xi()[3:6], xi_sub_mol) and bits_equal(off.xi()[6:9], xi_sub_mol)
    # 4. deterministic
    bis = small(oracle, levels=3)
    bis.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    bis.run(5)
    assert bits_equal(bis.xi(), xi)


@pytest.mark.parametrize("kw", [dict(smag_alpha=0.5, nu_sm_min=-0.1, eta_test=2.0),
                                dict(smag_alpha=0.3, nu_sm_min=-0.1, eta_test=3.0)])
def test_coarse_level_model_matches_numpy(oracle, kw):
    o = small(oracle)
    o.run(40)
    d = o.derived()
    g = o.grids[0]
    nx, ny = g.nx, g.ny
    f = g.f.reshape(ny, nx, 9).transpose(2, 0, 1).copy()
    fl = [a.reshape(ny, nx).copy() for a in (g.rho, g.ux, g.uy, g.temperature)]
    nu, alpha = d["nu_lu"], d["alpha_lu"]
    ref = DN.dynamic_sgs(f, *fl, tau=0.5 + 3.0 * nu, nu=nu, alpha_mol=alpha, **kw)
    o.set_dynamic_sgs(True, kw["smag_alpha"], kw["nu_sm_min"], kw["eta_test"])
    o.step()
    rel = lambda a, b: np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)  # noqa: E731
    assert rel(o.sgs_pre(0, 0), ref[0].ravel()) < 1e-10
    assert rel(o.sgs_pre(0, 1), ref[1].ravel()) < 1e-10
    st = o.dynamic_stats()
    assert abs(st[0] - ref[2].mean()) < 1e-12 and abs(st[2] - ref[3].mean()) < 1e-12
    assert st[1] == pytest.approx(ref[2].max(), abs=1e-12)


@pytest.mark.parametrize("levels", [2, 3])
def test_level_floors_fix_each_levels_viscosity(oracle, levels):
    """At nu = 0.0002 the tau floor binds on every level (4 alpha < (0.505-0.5)/3):
    nu_t >= (0.505-0.5)/3 - nu_level, alpha_t >= (0.505-0.5)/3 - alpha_level,
    reached on the quiet nodes, with nu_level = nu, 2 nu, 4 nu (and alpha
    likewise)."""
    o = small(oracle, levels=levels, nu_override=0.0002)
    o.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    o.run(3)
    d = o.derived()
    nu, alpha = d["nu_lu"], d["alpha_lu"]
    for gi, m in [(0, 1.0), (1, 2.0), (2, 2.0)] + ([(3, 4.0), (4, 4.0)] if levels == 3 else []):
        floor_nu = max(-0.1 * m * nu, MARGIN - m * nu)
        floor_al = max(-0.1 * m * alpha, MARGIN - m * alpha)
        assert floor_nu > 0.0 and floor_al > 0.0
        nt, at = o.sgs_pre(gi, 0), o.sgs_pre(gi, 1)
        assert nt.min() == floor_nu and np.count_nonzero(nt == floor_nu) > nt.size // 4, gi
        assert at.min() == floor_al, gi


def textbook_collide(f, g, rho, ux, uy, temp, nu, alpha, nu_t, alpha_t, beta):
    """TRT (even/odd parts, Lambda = 3/16) with the Guo source split by
    reflection, BGK for g; tau+ = 0.5 + 3 (nu + nu_t), tau_t = 0.5 + 3 (alpha + alpha_t)."""
    sh = lambda a: np.asarray(a, dtype=float)[:, None, None]  # noqa: E731
    f, g = sh(f), sh(g)
    r, u, v, t = (np.array([[x]]) for x in (rho, ux, uy, temp))
    fy = r * beta * t
    fneq = f - equilibrium(r, u, v)
    tp = 0.5 + 3.0 * (nu + nu_t)
    tm = 0.5 + (3.0 / 16.0) / (tp - 0.5)
    even, odd = 0.5 * (fneq + fneq[OPP]), 0.5 * (fneq - fneq[OPP])
    cu = CX * u + CY * v
    src = WW * ((CY - v) * fy / CS2 + cu * CY * fy / CS2 ** 2)
    s_odd, s_even = 0.5 * (src - src[OPP]), 0.5 * (src + src[OPP])
    fo = f - even / tp - odd / tm + (1.0 - 0.5 / tm) * s_odd + (1.0 - 0.5 / tp) * s_even
    tt = 0.5 + 3.0 * (alpha + alpha_t)
    go = g - (g - equilibrium(t, u, v)) / tt
    return fo.ravel(), go.ravel()


def test_per_node_collision_against_textbook(oracle):
    rng = np.random.default_rng(11)
    nu, alpha, beta = 0.004, 0.0056, 3e-5
    tau3 = oracle_relax(oracle, nu, alpha)
    for _ in range(50):
        rho = 1.0 + 0.02 * rng.uniform(-1, 1)
        ux, uy, t = 0.05 * rng.uniform(-1, 1), 0.05 * rng.uniform(-1, 1), rng.uniform(-0.5, 0.5)
        f = np.array([oracle.feq(i, rho, ux, uy) for i in range(9)]) * (1 + 0.01 * rng.uniform(-1, 1, 9))
        g = np.array([oracle.geq(i, t, ux, uy) for i in range(9)]) * (1 + 0.01 * rng.uniform(-1, 1, 9))
        nu_t, alpha_t = 0.01 * rng.uniform(), 0.01 * rng.uniform()
        fo, go, fy = oracle.oct_collide_node(f, g, rho, ux, uy, t, tau3, beta, pre=(nu_t, alpha_t))
        fr, gr = textbook_collide(f, g, rho, ux, uy, t, nu, alpha, nu_t, alpha_t, beta)
        assert np.max(np.abs(fo - fr)) < 1e-15 and np.max(np.abs(go - gr)) < 1e-15
        assert fy == rho * beta * t
        # nu_t = alpha_t = 0 is the level's fixed relaxation up to the rounding
        # of 0.5 + 3 (tau - 0.5) / 3 (a planted wrong tau would show at 1e-6)
        f0, g0, _ = oracle.oct_collide_node(f, g, rho, ux, uy, t, tau3, beta, pre=(0.0, 0.0))
        fx, gx, _ = oracle.oct_collide_node(f, g, rho, ux, uy, t, tau3, beta)
        assert np.max(np.abs(f0 - fx)) < 1e-16 and np.max(np.abs(g0 - gx)) < 1e-16
