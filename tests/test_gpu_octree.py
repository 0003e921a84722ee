"""Octree (multi-level) cavity on the GPU against the octree oracle.

Every level's populations, moments and Guo force must be bitwise equal to the
oracle's after any number of coarse steps (FP64, no contraction, IEEE
division, the reference's operand order on both sides); the diagnostics are
reduced on the host in the reference's order and are bitwise too.  The
reference's own octree tests (src/octree/tests.rs) are restated on the GPU
path.
"""
import numpy as np
import pytest

from cfd_cavity_amd import OctreeCavitySolver
from tests.helpers import bits_equal, first_mismatch

pytestmark = pytest.mark.gpu

FIELDS = (("rho", "rho"), ("ux", "ux"), ("uy", "uy"), ("temperature", "temperature"),
          ("fy_node", "fy"))


def pair(oracle, *args, **kw):
    g = OctreeCavitySolver(*args, **kw)
    o = oracle.OracleOctree(*args, **kw)
    return g, o


def assert_bitwise(g, o, what=""):
    assert g.n_grids == len(o.grids)
    for gi, lv in enumerate(o.grids):
        assert g.dims[gi] == (lv.nx, lv.ny)
        for name, oname in FIELDS:
            got, ref = g.get_field(gi, name), getattr(lv, oname)
            assert bits_equal(got, ref), f"{what} grid {gi} {name}: {first_mismatch(got, ref)}"
        for w in ("f", "g"):
            got, ref = g.get_pops(gi, w), getattr(lv, w)
            assert bits_equal(got, ref), f"{what} grid {gi} pops {w}: {first_mismatch(got, ref)}"


def assert_diag_equal(g, o):
    d = g.diagnostics()
    n = o.noise_stats()
    assert d["nusselt_hot"] == o.nusselt_hot()
    assert d["nusselt_cold"] == o.nusselt_cold()
    assert d["nusselt_hot_link0"] == o.nusselt_hot_link0()
    assert d["nusselt_hot_fine"] == o.nusselt_hot_fine(0, g.nx0 - 1.0)
    assert d["max_rho_dev"] == n["max_rho_dev"]
    assert d["max_rho_dev_ifaces"] == n["max_rho_dev_ifaces"]
    assert d["masse"] == n["masse"]
    assert d["max_u_all"] == o.max_u_all()
    assert d["guard"] == o.check_guards()
    assert d["time_step"] == o.time_step


@pytest.mark.parametrize("levels", [2, 3])
def test_initial_state_and_constants_bitwise(oracle, levels):
    g, o = pair(oracle, 32, 64, 4, 1.0e6, levels=levels)
    assert_bitwise(g, o, "init")
    d, od = g.derived(), o.derived()
    assert d["nu_lu"] == od["nu_lu"] and d["alpha_lu"] == od["alpha_lu"]
    assert d["beta_eff_coarse"] == od["beta_eff_coarse"]
    assert d["beta_eff_fine"] == od["beta_eff_fine"]
    assert d["spin_up_steps"] == od["spin_up_steps"]
    assert d["n_nodes_total"] == od["n_nodes_total"]
    for lvl, name in enumerate(("l0", "l1", "l2")):
        assert d["tau"][lvl] == [od[f"tau_plus_{name}"], od[f"tau_minus_{name}"],
                                 od[f"tau_t_{name}"]]
    assert d["xi_a"] == [od["xi_even_a"], od["xi_odd_a"], od["xi_thermal_a"]]
    assert d["xi_b"] == [od["xi_even_b"], od["xi_odd_b"], od["xi_thermal_b"]]


@pytest.mark.parametrize("args,kw,nsteps", [
    ((32, 64, 4, 1.0e6), dict(levels=2), 1),
    ((32, 64, 4, 1.0e6), dict(levels=2), 40),
    ((32, 64, 4, 1.0e6), dict(levels=3), 25),
    ((32, 64, 4, 1.0e6), dict(levels=3, odd_exact=True), 10),
    ((32, 64, 4, 1.0e6), dict(levels=2, nu_start_mult=3.0), 30),
    ((64, 256, 8, 1.0e6), dict(levels=2), 200),
    ((64, 256, 8, 1.0e7), dict(levels=3, seed=7), 60),
    ((32, 64, 4, 1.0e6), dict(levels=2, geq_order=1), 15),
    # a base grid too wide for the tiled sub-step kernel: it takes the
    # separate launches while the blocks take tiles
    ((128, 128, 8, 1.0e6), dict(levels=2), 20),
])
def test_steps_bitwise(oracle, args, kw, nsteps):
    g, o = pair(oracle, *args, **kw)
    g.run(nsteps)
    o.run(nsteps)
    assert g.time_step == nsteps
    assert_bitwise(g, o, f"{args} {kw} after {nsteps}")
    assert_diag_equal(g, o)


def test_single_steps_equal_graph_run(oracle):
    a = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=3)
    b = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=3)
    for _ in range(12):
        a.step_cavity()
        a.diagnostics()  # interleaved reads do not perturb the state
    b.run(12)
    for gi in range(a.n_grids):
        for w in ("f", "g"):
            assert bits_equal(a.get_pops(gi, w), b.get_pops(gi, w))


def test_random_state_step_bitwise(oracle):
    """A developed state with random non-equilibrium on every level, loaded
    into both, then stepped: every phase and transfer sees non-trivial data."""
    g, o = pair(oracle, 32, 64, 4, 1.0e6, levels=3)
    g.run(30)
    o.run(30)
    rng = np.random.default_rng(21)
    for gi, lv in enumerate(o.grids):
        n = lv.nx * lv.ny
        lv.f[:] = lv.f * (1.0 + 1e-3 * rng.uniform(-1, 1, (n, 9)))
        lv.g[:] = lv.g + 1e-4 * rng.uniform(-1, 1, (n, 9))
        lv.ux[:] = lv.ux + 1e-3 * rng.uniform(-1, 1, n)
        lv.fy[:] = lv.fy * (1.0 + 0.1 * rng.uniform(-1, 1, n))
        for w in ("f", "g"):
            g.set_pops(gi, w, getattr(lv, w))
        for name, oname in FIELDS:
            g.set_field(gi, name, getattr(lv, oname))
    assert_bitwise(g, o, "loaded")
    for _ in range(3):
        g.step_cavity()
        o.step()
    assert_bitwise(g, o, "random state, 3 steps")
    assert_diag_equal(g, o)


def test_set_from_snapshot_bitwise(oracle):
    g, o = pair(oracle, 32, 64, 4, 1.0e6, levels=3)
    g.run(20)
    o.run(20)
    fl = g.level_fields()
    rho, ux, uy, t = (np.concatenate([f[2 + q] for f in fl]) for q in range(4))
    rng = np.random.default_rng(4)
    ux = ux + 1e-3 * rng.uniform(-1, 1, ux.size)
    g.set_from_snapshot(rho, ux, uy, t)
    o.set_from_snapshot(rho, ux, uy, t)
    assert g.time_step == 0
    assert_bitwise(g, o, "after set_from_snapshot")
    g.run(15)
    o.run(15)
    assert_bitwise(g, o, "restart + 15")


# ---------------------------------------------------------------------------
# reference tests (src/octree/tests.rs) on the GPU path

def test_v2_mass_conserved_with_guo_force():                 # tests.rs:396-416
    s = OctreeCavitySolver(32, 64, 4, 1.0e6, 0.71, 0.02, 42, 1, False, 2, 1.0, None)
    for _ in range(10):
        s.step_cavity()
        assert abs(s.diagnostics()["masse"] - 1.0) < 1e-6


def test_v2_noise_stats_bounded():                           # tests.rs:418-447
    s = OctreeCavitySolver(32, 64, 4, 1.0e6)
    for step in range(20):
        s.step_cavity()
        n = s.noise_stats()
        assert n["max_rho_dev"] >= n["max_rho_dev_ifaces"]
        if step >= 5:
            assert n["max_rho_dev_ifaces"] < 1.0


def test_v2_nusselt_conductive_initial():                    # tests.rs:449-475
    d = OctreeCavitySolver(32, 64, 4, 1.0e6).diagnostics()
    assert abs(d["nusselt_hot"] - 1.0) < 0.05 and abs(d["nusselt_cold"] - 1.0) < 0.05


def test_v6_three_levels_stable_short():                     # tests.rs:548-575
    s = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=3)
    assert s.dims[3][0] == 7 and s.dims[4][0] == 7
    for _ in range(10):
        s.step_cavity()
        d = s.diagnostics()
        assert abs(d["masse"] - 1.0) < 1e-6
        assert d["max_rho_dev"] >= d["max_rho_dev_ifaces"] and d["max_rho_dev_ifaces"] < 1.0
        assert s.check_guards() is None


def test_errors():
    with pytest.raises(ValueError):
        OctreeCavitySolver(32, 64, 1, 1.0e6)           # solver.rs:356
    with pytest.raises(ValueError):
        OctreeCavitySolver(16, 64, 8, 1.0e6)           # solver.rs:357
    with pytest.raises(ValueError):
        OctreeCavitySolver(32, 64, 4, 1.0e6, levels=4)
    s = OctreeCavitySolver(32, 64, 4, 1.0e6)
    s.set_dynamic_sgs(True)            # buffers of zeros until the model first runs
    assert not np.any(s.sgs_pre(0)) and s.dynamic_stats() == (0.0,) * 5
    s.set_dynamic_sgs(False)


# ---------------------------------------------------------------------------
# OCTSNAP1 restart (src/octree/tests.rs:795-965)

def test_snapshot_restart_restores_distributions(oracle, tmp_path):   # tests.rs:837-945
    sim = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=3)
    sim.run(5)
    path = tmp_path / "snap.bin"
    sim.write_snapshot(path, 5)
    restored = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=3)
    assert restored.restart(path) == 5
    assert restored.time_step == 0
    for gi, (nx, ny, rho, ux, uy, t) in enumerate(sim.level_fields()):
        assert bits_equal(restored.get_field(gi, "rho"), rho)
        assert bits_equal(restored.get_field(gi, "temperature"), t)
        f = restored.get_pops(gi, "f")
        k = (ny // 2) * nx + nx // 2  # an interior node: its velocity is the exported one
        assert bits_equal(restored.get_field(gi, "ux")[k], ux[k])
        assert bits_equal(f[k], np.array([oracle.feq(i, rho[k], ux[k], uy[k]) for i in range(9)]))
    sim.step_cavity()
    restored.step_cavity()
    a, b = sim.diagnostics(), restored.diagnostics()
    assert abs(a["masse"] - b["masse"]) < 5e-6
    assert abs(a["max_rho_dev"] - b["max_rho_dev"]) < 5e-3
    assert abs(a["nusselt_hot"] - b["nusselt_hot"]) < 5e-2
    ra, rb = sim.get_field(0, "rho"), restored.get_field(0, "rho")
    assert np.max(np.abs(ra - rb)) < 5e-4


def test_snapshot_restart_layout_mismatch_is_error(tmp_path):        # tests.rs:947-965
    sim = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=3)
    sim.run(2)
    path = tmp_path / "snap.bin"
    sim.write_snapshot(path, 2)
    other = OctreeCavitySolver(32, 64, 6, 1.0e6, levels=3)
    with pytest.raises(ValueError, match="layout"):
        other.restart(path)
    two = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=2)
    with pytest.raises(ValueError, match="incompatible"):
        two.restart(path)


# ---------------------------------------------------------------------------
# Dynamic mixed-scale SGS (src/octree/smag.rs, solver.rs:647-947).
#
# Tolerance.  The model's mixed forms use the device's pow (the correctly
# rounded square root for alpha = 0.5) where the oracle, like the reference,
# uses the C library's pow, and the levels' mean nu_t / alpha_t that set the
# interface factors are tree sums on the device where the reference sums
# sequentially.  Both differ by a few ulp, so: nu_t / alpha_t per node within
# 1e-13 of the level's largest value and the factors within 1e-14 after one
# step; after 20 steps populations and fields within 1e-10 (relative to their
# largest value; 1e-9 at nu = 0.0005), statistics within 1e-9, the factors
# within 1e-12 and nu_t /
# alpha_t per node within 1e-6 (the Lilly ratio num / den of a box where den
# is small amplifies the state's round-off differences).  The floors, which
# pow does not reach, are bitwise.

def relmax(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def assert_close(g, o, tol, what):
    for gi, lv in enumerate(o.grids):
        for name, oname in FIELDS:
            r = relmax(g.get_field(gi, name), getattr(lv, oname))
            assert r < tol, (what, gi, name, r)
        for w in ("f", "g"):
            r = relmax(g.get_pops(gi, w), getattr(lv, w))
            assert r < tol, (what, gi, w, r)


def assert_model_close(g, o, tol_pre, tol_stats, what, tol_xi=1e-14):
    for gi in range(len(o.grids)):
        for which in (0, 1):
            r = relmax(g.sgs_pre(gi, which), o.sgs_pre(gi, which))
            assert r < tol_pre, (what, gi, which, r)
    for a, b in zip(g.dynamic_stats(), o.dynamic_stats()):
        assert abs(a - b) <= tol_stats * max(abs(b), 1e-3), (what, g.dynamic_stats(),
                                                             o.dynamic_stats())
    xg, xo = g.interface_factors(), o.xi()
    assert np.max(np.abs(xg - xo)) < tol_xi, (what, xg, xo)


@pytest.mark.parametrize("kw,model,tol20", [
    (dict(levels=2), (0.5, -0.1, 2.0), 1e-10),
    (dict(levels=3), (0.5, -0.1, 2.0), 1e-10),
    (dict(levels=3, odd_exact=True), (0.3, -0.1, 3.0), 1e-10),
    # test-filter width 1 (tiled model), and width 5: the three passes, whose
    # coefficient pass writes the workgroup sums the factor update reads
    (dict(levels=2), (0.5, -0.1, 1.0), 1e-10),
    (dict(levels=3), (0.5, -0.1, 5.0), 1e-10),
    (dict(levels=2, nu_start_mult=3.0), (0.5, -0.1, 2.0), 1e-10),  # direct launches, ramped nu
    # the tau floor binds; at nu = 0.0005 the differences grow to ~2e-10 of
    # the strips' small velocities in 20 steps
    (dict(levels=3, nu_override=0.0005), (0.7, -0.5, 2.4), 1e-9),
])
def test_dynamic_matches_oracle(oracle, kw, model, tol20):
    g, o = pair(oracle, 32, 64, 4, 1.0e6, **kw)
    g.run(30)
    o.run(30)
    assert_bitwise(g, o, "before the model")
    g.set_dynamic_sgs(True, *model)
    o.set_dynamic_sgs(True, *model)
    g.step_cavity()
    o.step()
    assert_model_close(g, o, 1e-13, 1e-12, "1 step")
    assert_close(g, o, 1e-13, "1 step")
    g.run(19)
    o.run(19)
    assert_close(g, o, tol20, "20 steps")
    assert_model_close(g, o, 1e-6, 1e-9, "20 steps", tol_xi=1e-12)


def test_dynamic_64x256_bl8(oracle):
    g, o = pair(oracle, 64, 256, 8, 1.0e6, levels=2)
    g.run(100)
    o.run(100)
    g.set_dynamic_sgs(True)
    o.set_dynamic_sgs(True)
    g.run(40)
    o.run(40)
    assert_close(g, o, 1e-10, "64x256, 40 dynamic steps")
    d, od = g.diagnostics(), o.nusselt_hot()
    assert abs(d["nusselt_hot"] - od) < 1e-10 * abs(od)


def test_dynamic_switched_off_keeps_its_factors(oracle):
    """After the model is switched off the collisions use the levels' fixed
    relaxation times again and the interfaces keep the last recomputed
    factors (solver.rs:690-722: xi is only rewritten with the model on)."""
    g, o = pair(oracle, 32, 64, 4, 1.0e6, levels=3)
    for s in (g, o):
        s.set_dynamic_sgs(True)
    g.run(5)
    o.run(5)
    x5 = g.interface_factors()
    for s in (g, o):
        s.set_dynamic_sgs(False)
    g.run(10)
    o.run(10)
    assert bits_equal(g.interface_factors(), x5)
    assert_close(g, o, 1e-10, "model off after 5 dynamic steps")


def test_dynamic_is_deterministic_and_follows_effective_taus():   # tests.rs:686-786
    runs = []
    for _ in range(2):
        s = OctreeCavitySolver(32, 64, 4, 1.0e6, levels=3)
        mol = s.interface_factors()
        s.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
        s.run(5)
        runs.append(s)
    a, b = runs
    assert bits_equal(a.interface_factors(), b.interface_factors())
    for gi in range(a.n_grids):
        assert bits_equal(a.get_pops(gi, "f"), b.get_pops(gi, "f"))
        assert bits_equal(a.sgs_pre(gi), b.sgs_pre(gi))
    xi = a.interface_factors()
    assert xi[0] != mol[0] and xi[2] != mol[2] and xi[3] != mol[3]
    assert abs(xi[0] - mol[0]) / mol[0] < 0.1
    cs_mean, cs_max, k_mean, k_max, nu_t_max = a.dynamic_stats()
    assert 0.0 <= cs_max <= 0.04 and cs_mean <= cs_max + 1e-12 and 0.0 <= k_max <= 0.04
    assert nu_t_max > 0.0 and a.check_guards() is None


@pytest.mark.parametrize("levels", [2, 3])
def test_dynamic_floors_per_level_bitwise(oracle, levels):
    g, o = pair(oracle, 32, 64, 4, 1.0e6, levels=levels, nu_override=0.0002)
    for s in (g, o):
        s.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    g.run(3)
    o.run(3)
    for gi in range(len(o.grids)):
        for which in (0, 1):
            got, ref = g.sgs_pre(gi, which), o.sgs_pre(gi, which)
            floor = ref.min()
            assert got.min() == floor, (gi, which)
            assert np.count_nonzero((got == floor) != (ref == floor)) <= 2, (gi, which)
