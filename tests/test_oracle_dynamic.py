"""Pins of the oracle's dynamic mixed-scale SGS model (no GPU).

The oracle restates compute_dynamic_sgs (src/uniform/solver.rs:562-828) node
by node.  It is pinned by
  - the reference's own tests of the model (src/uniform/tests.rs:516-578);
  - closed forms: a fluid at rest, and a test filter of width one (eta = 1),
    both make every Lilly numerator vanish, so nu_t and alpha_t sit exactly on
    their floors;
  - an independent vectorised numpy formulation (tests/dynamic_sgs_numpy.py)
    for the model alone and inside full steps of the numpy textbook solver,
    over several (alpha, eta, nu_sm_min) and with the tau floor active;
  - a planted mistake in the numpy formulation that the comparison detects.
"""
import numpy as np
import pytest

import tests.dynamic_sgs_numpy as DN
from tests.helpers import SMALL, args_of, bits_equal
from tests.test_oracle_pins import compare_textbook, textbook_of

EPS = 1e-12  # tests.rs uses EPS = 1e-12


def small(oracle, **kw):
    return oracle.OracleSolver(*args_of(SMALL), 1, **kw)


def low_nu(nu_override):
    """Solver options for a run below the calibrated viscosity (the reference's
    Ra = 1e10 setting: --nu with alpha_mult 2.5 keeps tau_t above 0.505)."""
    return {} if nu_override is None else dict(nu_override=nu_override, alpha_mult=2.5)


def floors(o, nu_sm_min):
    d = o.derived()
    margin = (0.505 - 0.5) / 3.0
    return (max(nu_sm_min * d["nu"], margin - d["nu"]),
            max(nu_sm_min * d["alpha"], margin - d["alpha"]))


# ---------------------------------------------------------------------------
# the reference's tests (src/uniform/tests.rs)

def test_lilly_coefficient_clamps(oracle):                 # tests.rs:516-529
    L = oracle.lilly_coefficient
    assert L(1.0, 0.0) == 0.0
    assert L(1.0, -1e-12) == 0.0
    assert L(float("nan"), 1.0) == 0.0
    assert abs(L(0.02, 1.0) - 0.01) < EPS
    assert L(100.0, 1.0) == 0.04
    assert L(-100.0, 1.0) == 0.0
    assert L(1.0, float("inf")) == 0.0 and L(float("inf"), 1.0) == 0.0


def test_mixed_factor_geometric_mean(oracle):              # tests.rs:531-541
    m = oracle.mixed_factor
    a, b = 4.0, 9.0
    assert abs(m(a, b, 0.5) - np.sqrt(a * b)) < 1e-12
    assert abs(m(a, b, 0.0) - b) < 1e-12
    assert abs(m(a, b, 1.0) - a) < 1e-12


def test_dynamic_off_matches_fixed_cs(oracle):              # tests.rs:543-558
    s = small(oracle)
    s.set_dynamic_sgs(False, 0.5, -0.1, 2.0)
    r = small(oracle)
    s.run(30)
    r.run(30)
    assert s.sgs_pre(0) is None and s.sgs_pre(1) is None
    assert bits_equal(s.nu_t, r.nu_t)
    assert bits_equal(s.f, r.f) and bits_equal(s.g, r.g)


def test_dynamic_on_runs_sane(oracle):                      # tests.rs:560-578
    s = small(oracle)
    s.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    s.run(5)
    assert s.sgs_pre(0) is not None and s.sgs_pre(1) is not None
    mean, mx, kmean, kmx = s.dynamic_stats()
    assert np.isfinite(mean) and np.isfinite(mx)
    assert 0.0 <= mx <= 0.04
    assert mean <= mx + EPS
    assert 0.0 <= kmx <= 0.04 and kmean <= kmx + EPS
    assert np.all(np.isfinite(s.nu_t))


# ---------------------------------------------------------------------------
# closed forms

@pytest.mark.parametrize("nu_override", [None, 0.0005])
def test_fluid_at_rest_sits_on_the_floors(oracle, nu_override):
    """u = 0 and f = feq: pi_neq, S~, S-, the similarity scales, L and M all
    vanish, so C = 0 and nu_t = max(C nu_mix, floor) is exactly +0 or, when
    positive, the floor max(nu_sm_min * nu, (0.505 - 0.5)/3 - nu); the same
    for alpha_t, whatever T is."""
    s = small(oracle, **low_nu(nu_override))
    n = s.nx * s.ny
    rng = np.random.default_rng(5)
    s.init_from_fields(1.0 + 0.01 * rng.uniform(-1, 1, n), np.zeros(n), np.zeros(n),
                       rng.uniform(-0.5, 0.5, n))
    s.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    s.compute_dynamic_sgs()
    nf, af = floors(s, -0.1)
    assert bits_equal(s.sgs_pre(0), np.full(n, max(nf, 0.0)))
    assert bits_equal(s.sgs_pre(1), np.full(n, max(af, 0.0)))
    assert s.dynamic_stats() == (0.0, 0.0, 0.0, 0.0)
    if nu_override is not None:
        assert nf > 0.0  # the tau floor is the active bound for a small nu


def test_unit_test_filter_sits_on_the_floors(oracle):
    """eta = 1: the box filter is the identity, so L_ij = u_i u_j - u_i u_j = 0
    and L_j = 0 exactly; the Lilly numerators vanish on a developed flow too."""
    s = small(oracle)
    s.run(200)
    s.set_dynamic_sgs(True, 0.5, -0.1, 1.0)
    s.compute_dynamic_sgs()
    nf, af = floors(s, -0.1)
    n = s.nx * s.ny
    assert bits_equal(s.sgs_pre(0), np.full(n, max(nf, 0.0)))
    assert bits_equal(s.sgs_pre(1), np.full(n, max(af, 0.0)))
    assert s.dynamic_stats() == (0.0, 0.0, 0.0, 0.0)


# ---------------------------------------------------------------------------
# independent numpy formulation

def numpy_model(o, **kw):
    nx, ny = o.nx, o.ny
    d = o.derived()
    step = o.time_step
    nu_eff = d["nu"]
    if step < d["spin_up_steps"]:
        x = step / d["spin_up_steps"]
        start = d["nu"] * d["nu_start_mult"]
        nu_eff = start * (d["nu"] / start) ** x
    tau = 0.5 + 3.0 * max(nu_eff, 0.002 / 3.0)
    f = o.f.reshape(ny, nx, 9).transpose(2, 0, 1)

    def fl(a):
        return a.reshape(ny, nx)

    return DN.dynamic_sgs(f, fl(o.rho), fl(o.ux), fl(o.uy), fl(o.temperature), tau=tau,
                          nu=d["nu"], alpha_mol=d["alpha"], **kw)


def relmax(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


MODELS = [dict(smag_alpha=0.5, nu_sm_min=-0.1, eta_test=2.0),
          dict(smag_alpha=0.3, nu_sm_min=-0.1, eta_test=3.0),
          dict(smag_alpha=0.7, nu_sm_min=-0.5, eta_test=2.4),
          dict(smag_alpha=1.0, nu_sm_min=-0.1, eta_test=2.0),
          dict(smag_alpha=0.0, nu_sm_min=-0.1, eta_test=4.0)]


@pytest.mark.parametrize("kw", MODELS)
@pytest.mark.parametrize("nu_override", [None, 0.0005])
def test_oracle_model_matches_numpy(oracle, kw, nu_override):
    s = small(oracle, **low_nu(nu_override))
    s.run(300)
    ref = numpy_model(s, **kw)
    s.set_dynamic_sgs(True, kw["smag_alpha"], kw["nu_sm_min"], kw["eta_test"])
    s.compute_dynamic_sgs()
    nu_t, al_t = s.sgs_pre(0), s.sgs_pre(1)
    assert relmax(nu_t, ref[0].ravel()) < 1e-10
    assert relmax(al_t, ref[1].ravel()) < 1e-10
    mean, mx, kmean, kmx = s.dynamic_stats()
    assert abs(mean - ref[2].mean()) < 1e-12 and abs(kmean - ref[3].mean()) < 1e-12
    assert mx == pytest.approx(ref[2].max(), abs=1e-12)
    assert kmx == pytest.approx(ref[3].max(), abs=1e-12)
    # the model is not degenerate here: some coefficients are strictly inside
    # (0, 0.04), and the floors bind somewhere only when nu is small
    inside = (ref[2] > 0) & (ref[2] < 0.04)
    assert inside.mean() > 0.05
    nf, _ = floors(s, kw["nu_sm_min"])
    assert (np.any(nu_t == nf)) == (nu_override is not None)


def test_numpy_comparison_detects_planted_mistake(oracle, monkeypatch):
    """Plausible mistakes planted in the numpy formulation (a 2-D trace L_kk/2
    in place of the reference's 3-D L_kk/3; d/dy in place of d/dx; a box
    window shifted by one node) move the model far beyond the tolerance of the
    comparisons above."""
    s = small(oracle)
    s.run(300)
    s.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    s.compute_dynamic_sgs()
    got = s.sgs_pre(0)
    assert relmax(got, numpy_model(s)[0].ravel()) < 1e-10
    box = DN.box
    for name, value in (("TRACE_DIM", 2.0), ("d_dx", DN.d_dy),
                        ("box", lambda a, w: np.roll(box(a, w), 1, axis=1))):
        with monkeypatch.context() as m:
            m.setattr(DN, name, value)
            bad = numpy_model(s)[0].ravel()
        assert relmax(got, bad) > 1e-4, name


@pytest.mark.parametrize("kw,nsteps", [(MODELS[0], 25), (MODELS[2], 15)])
def test_oracle_dynamic_steps_match_textbook(oracle, kw, nsteps):
    """Full steps with the dynamic model: the oracle against the numpy textbook
    solver driving the numpy model, from a developed state."""
    s = small(oracle)
    s.run(250)
    tb = textbook_of(s, dynamic=kw)
    tb.time_step = s.time_step
    s.set_dynamic_sgs(True, kw["smag_alpha"], kw["nu_sm_min"], kw["eta_test"])
    for _ in range(nsteps):
        s.step()
        tb.step()
    compare_textbook(s, tb, 1e-12)


def test_values_persist_after_switching_off(oracle):
    """solver.rs keeps nu_t_pre / alpha_t_pre once computed; collide keeps
    using them after set_dynamic_sgs(false): the state then departs from a
    fixed-Cs run, and the stored values stay those of the last dynamic step."""
    s = small(oracle)
    s.run(100)
    s.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    s.run(3)
    pre = s.sgs_pre(0)
    stats = s.dynamic_stats()
    s.set_dynamic_sgs(False, 0.5, -0.1, 2.0)
    s.run(2)
    assert bits_equal(s.sgs_pre(0), pre) and s.dynamic_stats() == stats
    assert bits_equal(s.nu_t, pre)  # collide reports the value it used
