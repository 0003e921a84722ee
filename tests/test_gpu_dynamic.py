"""Dynamic mixed-scale SGS on the GPU (compute_dynamic_sgs,
src/uniform/solver.rs:562-828) against the oracle.

Tolerance.  The model's mixed forms a^alpha b^(1-alpha) are evaluated with the
C library's pow in the reference (Rust powf) and in the oracle; pow is not
correctly rounded, and the device's pow (or, for alpha = 0.5, its correctly
rounded square root) differs from it by at most a few ulp on a small fraction
of nodes.  Every other operation keeps the reference's order with IEEE
results.  Hence: nu_t / alpha_t per node within 1e-13 of the oracle's largest
value after one step, fields within 1e-10 (relative to their largest value)
after 30 steps, coefficient statistics within 1e-9.  Cases where pow is not
reached (a fluid at rest, a unit test filter) and the model-off path are
bitwise.
"""
import numpy as np
import pytest

from cfd_cavity_amd import CavityBatch, CavitySolver
from cfd_cavity_amd._lib import CavlbmError, E_UNSUPPORTED
from cfd_cavity_amd.adjoint import nu_window
from tests.helpers import SMALL, args_of, bits_equal, first_mismatch
from tests.test_gpu_parity import FIELDS, assert_state_bitwise, pair

pytestmark = pytest.mark.gpu


def low_nu(nu_override):
    return {} if nu_override is None else dict(nu_override=nu_override, alpha_mult=2.5)


def relmax(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def assert_state_close(g, o, tol, what=""):
    for name in FIELDS:
        got, ref = g.get_field(name), getattr(o, name)
        assert relmax(got, ref) < tol, (what, name, relmax(got, ref))
    for w in ("f", "g"):
        assert relmax(g.get_pops(w), getattr(o, w)) < tol, (what, w)


def assert_stats_close(g, o, tol):
    sg, so = g.dynamic_stats(), o.dynamic_stats()
    for a, b in zip(sg, so):
        assert abs(a - b) <= tol * max(abs(b), 1e-3), (sg, so)


def test_model_off_is_the_fixed_cs_path(oracle):          # tests.rs:543-558
    g, o = pair(oracle)
    g.set_dynamic_sgs(False, 0.5, -0.1, 2.0)
    g.run(30)
    o.run(30)
    assert g.nu_t_pre is None and g.alpha_t_pre is None
    assert g.dynamic_stats() == (0.0, 0.0, 0.0, 0.0)
    assert_state_bitwise(g, o, "model off")


def test_model_on_runs_sane():                            # tests.rs:560-578
    g = CavitySolver(*args_of(SMALL))
    g.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    g.run(5)
    assert g.nu_t_pre is not None and g.alpha_t_pre is not None
    mean, mx, kmean, kmx = g.dynamic_stats()
    assert np.isfinite(mean) and np.isfinite(mx)
    assert 0.0 <= mx <= 0.04 and mean <= mx + 1e-12
    assert 0.0 <= kmx <= 0.04 and kmean <= kmx + 1e-12
    assert np.all(np.isfinite(g.get_field("nu_t")))


MODELS = [(0.5, -0.1, 2.0), (0.3, -0.1, 3.0), (0.7, -0.5, 2.4)]


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("nu_override", [None, 0.0005])
def test_model_matches_oracle(oracle, model, nu_override):
    g, o = pair(oracle, **low_nu(nu_override))
    g.run(300)
    o.run(300)
    assert_state_bitwise(g, o, "before the model")
    g.set_dynamic_sgs(True, *model)
    o.set_dynamic_sgs(True, *model)
    g.step()
    o.step()
    assert relmax(g.nu_t_pre, o.sgs_pre(0)) < 1e-13
    assert relmax(g.alpha_t_pre, o.sgs_pre(1)) < 1e-13
    assert bits_equal(g.get_field("nu_t"), g.nu_t_pre)   # collide used these values
    assert_stats_close(g, o, 1e-12)
    assert_state_close(g, o, 1e-13, "1 step")
    g.run(29)
    o.run(29)
    assert_state_close(g, o, 1e-10, "30 steps")
    assert_stats_close(g, o, 1e-9)


def test_model_matches_oracle_64x256_ra1e6(oracle):
    cfg = dict(SMALL, nx=64, ny=256, ra=1.0e6)
    g, o = pair(oracle, cfg, oracle_jobs=8)
    g.run(500)
    o.run(500)
    g.set_dynamic_sgs(True)
    o.set_dynamic_sgs(True)
    g.run(100)
    o.run(100)
    assert_state_close(g, o, 1e-10, "64x256 Ra1e6, 100 dynamic steps")
    assert_stats_close(g, o, 1e-9)
    dg, do = g.diagnostics(), o.diagnostics()
    assert abs(dg["nusselt_hot"] - do["nusselt_hot"]) < 1e-10 * abs(do["nusselt_hot"])


@pytest.mark.parametrize("nu_override", [None, 0.0005])
def test_fluid_at_rest_is_exact(oracle, nu_override):
    """u = 0, f = feq: every Lilly numerator is zero, nu_t / alpha_t are +0 or
    the positive floor, bitwise as in the oracle."""
    g, o = pair(oracle, **low_nu(nu_override))
    n = SMALL["nx"] * SMALL["ny"]
    rng = np.random.default_rng(5)
    fields = (1.0 + 0.01 * rng.uniform(-1, 1, n), np.zeros(n), np.zeros(n),
              rng.uniform(-0.5, 0.5, n))
    g.init_from_fields(*fields)
    o.init_from_fields(*fields)
    for s in (g, o):
        s.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
        s.step()
    assert bits_equal(g.nu_t_pre, o.sgs_pre(0)) and bits_equal(g.alpha_t_pre, o.sgs_pre(1))
    assert g.dynamic_stats() == (0.0, 0.0, 0.0, 0.0)
    assert_state_bitwise(g, o, "rest, 1 step")


def test_unit_test_filter_is_exact(oracle):
    g, o = pair(oracle)
    g.run(200)
    o.run(200)
    for s in (g, o):
        s.set_dynamic_sgs(True, 0.5, -0.1, 1.0)
        s.step()
    assert bits_equal(g.nu_t_pre, o.sgs_pre(0)), first_mismatch(g.nu_t_pre, o.sgs_pre(0))
    assert bits_equal(g.alpha_t_pre, o.sgs_pre(1))
    assert_state_bitwise(g, o, "eta = 1")


def test_values_persist_after_switching_off(oracle):
    g, o = pair(oracle)
    g.run(100)
    o.run(100)
    for s in (g, o):
        s.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    g.run(3)
    o.run(3)
    pre, stats = g.nu_t_pre, g.dynamic_stats()
    for s in (g, o):
        s.set_dynamic_sgs(False, 0.5, -0.1, 2.0)
    g.run(150)  # past the spin-up (200): fused steps, then graph replays of them
    o.run(150)
    assert bits_equal(g.nu_t_pre, pre) and g.dynamic_stats() == stats
    assert bits_equal(g.get_field("nu_t"), pre)
    assert_state_close(g, o, 1e-10, "after switching off")


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_kept_values_fused_steps_equal_mform_steps(storage):
    """With the model switched off and its values kept, steps from the C-form
    collide with them in the fused step (MODE_FUSED_PRE, graph-replayed after
    the spin-up); a solver made to start every step from the M-form (a field
    written back after each step) collides with them there (MODE_FIRST_PRE).
    Both give bitwise the same state."""
    a, b = (CavitySolver(*args_of(SMALL), 1, storage=storage) for _ in range(2))
    for s in (a, b):
        s.run(120)
        s.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
        s.run(4)
        s.set_dynamic_sgs(False, 0.5, -0.1, 2.0)
    a.run(130)
    for _ in range(130):
        b.set_field("rho", b.get_field("rho"))
        b.run(1)
    assert a.time_step == b.time_step == 254
    for name in FIELDS:
        assert bits_equal(a.get_field(name), b.get_field(name)), name
    for w in ("f", "g"):
        assert bits_equal(a.get_pops(w), b.get_pops(w)), w


def test_fp32_storage_follows_fp64():
    runs = {}
    for storage in ("fp64", "fp32"):
        s = CavitySolver(*args_of(dict(SMALL, nx=32, ny=128, ra=1e6)), 1, storage=storage)
        s.run(400)
        s.set_dynamic_sgs(True)
        s.run(200)
        runs[storage] = (s.diagnostics(), s.dynamic_stats())
    d64, d32 = runs["fp64"][0], runs["fp32"][0]
    assert d32["all_finite"]
    assert abs(d32["nusselt_hot"] - d64["nusselt_hot"]) < 1e-4 * abs(d64["nusselt_hot"])
    assert abs(runs["fp32"][1][0] - runs["fp64"][1][0]) < 1e-3


def test_unsupported_combinations():
    b = CavityBatch([dict(nx=16, ny=40, ra=1e5, cs=0.1)], storage="fp64")
    with pytest.raises(CavlbmError) as e:
        b[0].set_dynamic_sgs(True)
    assert e.value.code == E_UNSUPPORTED
    b[0].set_dynamic_sgs(False)
    b.close()
    g = CavitySolver(*args_of(SMALL))
    g.set_dynamic_sgs(True)
    g.run(2)
    with pytest.raises(CavlbmError) as e:
        nu_window(g, None, 0.1, 10)
    assert e.value.code == E_UNSUPPORTED
    with pytest.raises(CavlbmError) as e:
        g.phase("collide", g.time_step)
    assert e.value.code == E_UNSUPPORTED


def test_graph_replayed_dynamic_steps_equal_single_steps():
    """After the spin-up, runs of dynamic steps replay captured graphs (48
    steps each); they must equal the same steps issued one by one."""
    runs = []
    for single in (False, True):
        s = CavitySolver(*args_of(dict(SMALL, nx=32, ny=128, ra=1e6)))
        s.run(250)
        s.set_dynamic_sgs(True, 0.3, -0.1, 3.0)
        if single:
            for _ in range(130):
                s.step()
        else:
            s.run(130)
        runs.append(s)
    a, b = runs
    for w in ("f", "g"):
        assert bits_equal(a.get_pops(w), b.get_pops(w))
    assert bits_equal(a.nu_t_pre, b.nu_t_pre) and a.dynamic_stats() == b.dynamic_stats()


def test_reads_between_dynamic_steps_do_not_perturb():
    """A dynamic step from the C-form materializes only the f planes and
    fields, and the collision runs the coefficient stage itself; a read in
    between (which materializes the whole M-form, so the next step skips its
    own materialization) must not change the trajectory, the per-node values
    or the statistics."""
    runs = []
    for reads in (False, True):
        s = CavitySolver(*args_of(dict(SMALL, nx=32, ny=128, ra=1e6)))
        s.run(120)
        s.set_dynamic_sgs(True)
        for chunk in (1, 2, 5, 60, 3):
            s.run(chunk)
            if reads:
                s.get_field("temperature")
                s.get_pops("g")
        runs.append(s)
    a, b = runs
    for name in FIELDS:
        assert bits_equal(a.get_field(name), b.get_field(name)), name
    for w in ("f", "g"):
        assert bits_equal(a.get_pops(w), b.get_pops(w)), w
    assert bits_equal(a.nu_t_pre, b.nu_t_pre) and a.dynamic_stats() == b.dynamic_stats()


def _dyn_pair(cfg, model, spin):
    runs = []
    for one in (True, False):
        s = CavitySolver(*args_of(cfg))
        s.run(spin)
        s.set_dynamic_fusion(one)
        s.set_dynamic_sgs(True, *model)
        runs.append(s)
    return runs


def assert_same_dynamic_state(a, b, what):
    for name in FIELDS:
        assert bits_equal(a.get_field(name), b.get_field(name)), (what, name)
    for w in ("f", "g"):
        ga, gb = a.get_pops(w), b.get_pops(w)
        assert bits_equal(ga, gb), (what, w, first_mismatch(ga, gb))
    assert bits_equal(a.nu_t_pre, b.nu_t_pre), (what, "nu_t_pre")
    assert bits_equal(a.alpha_t_pre, b.alpha_t_pre), (what, "alpha_t_pre")
    assert a.dynamic_stats() == b.dynamic_stats(), what


@pytest.mark.parametrize("model", [(0.5, -0.1, 2.0), (0.5, -0.1, 1.0), (0.3, -0.1, 3.0),
                                   (0.7, -0.5, 2.4)])
@pytest.mark.parametrize("cfg", [SMALL, dict(SMALL, nx=50, ny=70, ra=1e6),
                                 dict(SMALL, nx=64, ny=256, ra=1e6)])
def test_one_launch_step_equals_four_launches(cfg, model):
    """The one-launch dynamic step (tiles of the model's stencil through LDS)
    gives the four-launch step's state bit for bit: test-filter widths 1, 2, 3,
    lattices narrower than a tile and not multiples of it, direct steps, reads
    in between and graph replays (48 steps)."""
    a, b = _dyn_pair(cfg, model, 190)   # steps 190..199 in the buoyancy ramp
    assert a.dynamic_launches == 1 and b.dynamic_launches == 4
    for chunk in (1, 2, 7, 60, 3):
        a.run(chunk)
        b.run(chunk)
        assert_same_dynamic_state(a, b, f"{cfg['nx']}x{cfg['ny']} {model} t={a.time_step}")


def test_one_launch_step_matches_oracle(oracle):
    g, o = pair(oracle)
    g.run(200)
    o.run(200)
    g.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    o.set_dynamic_sgs(True, 0.5, -0.1, 2.0)
    assert g.dynamic_launches == 1
    g.step()
    o.step()
    assert relmax(g.nu_t_pre, o.sgs_pre(0)) < 1e-13
    assert_state_close(g, o, 1e-13, "1 step")
    g.run(59)
    o.run(59)
    assert_state_close(g, o, 1e-10, "60 steps")
    assert_stats_close(g, o, 1e-9)


def test_wide_filter_keeps_four_launches():
    s = CavitySolver(*args_of(SMALL))
    s.set_dynamic_sgs(True, 0.5, -0.1, 5.0)      # width 5: beyond the one-launch tiles
    assert s.dynamic_launches == 4
    s.run(3)
    assert s.diagnostics()["all_finite"]
    f32 = CavitySolver(*args_of(SMALL), 1, storage="fp32")
    f32.set_dynamic_sgs(True, 0.5, -0.1, 5.0)
    assert f32.dynamic_launches == 4


@pytest.mark.parametrize("storage", ["fp32", "fp16_offset"])
@pytest.mark.parametrize("model", [(0.5, -0.1, 2.0), (0.5, -0.1, 1.0), (0.3, -0.1, 3.0)])
@pytest.mark.parametrize("nx,ny", [(50, 70), (3, 40), (33, 17)])
def test_low_precision_one_launch_equals_four(storage, model, nx, ny):
    """FP32 / FP16-offset storage: the one-launch dynamic step (the pull and
    collision in the storage's arithmetic, the model in FP64 on the pulled
    values) gives the four-launch step's state bit for bit, through direct
    steps and graph replays."""
    runs = []
    for fuse in (True, False):
        s = CavitySolver(nx, ny, 1e6, 0.71, 0.1, 0.9, 0.02, 42, 1, storage=storage)
        s.run(190)
        s.set_dynamic_fusion(fuse)
        s.set_dynamic_sgs(True, *model)
        assert s.dynamic_launches == (1 if fuse else 4)
        runs.append(s)
    for chunk in (1, 2, 7, 60):
        for s in runs:
            s.run(chunk)
        assert_same_dynamic_state(runs[0], runs[1], f"{storage} {model} t={runs[0].time_step}")


@pytest.mark.parametrize("nx,ny,eta,geq,hot", [
    (3, 40, 2.0, 2, False),      # the narrowest lattice: every pulled node is a wall or next to one
    (17, 3, 3.0, 2, False),      # three rows
    (5, 7, 1.0, 1, False),       # linear g equilibrium, unit filter
    (40, 33, 2.0, 2, True),      # hot-wall profile, 33 rows: a last tile of one row
    (33, 130, 3.0, 1, True),
])
def test_one_launch_step_edge_lattices(nx, ny, eta, geq, hot):
    """The one-launch dynamic step on lattices that exercise its clamped
    windows and partial tiles, with the linear g equilibrium and a hot-wall
    profile: bitwise the four-launch step."""
    runs = []
    for one in (True, False):
        s = CavitySolver(nx, ny, 1e5, 0.71, 0.1, 0.9, 0.02, 42, 1, geq_order=geq)
        if hot:
            s.set_hot_wall(0.5 + 0.05 * np.sin(np.linspace(0.0, 3.0, ny)))
        s.run(30)
        s.set_dynamic_fusion(one)
        s.set_dynamic_sgs(True, 0.5, -0.1, eta)
        assert s.dynamic_launches == (1 if one else 4)
        s.run(25)
        runs.append(s)
    assert_same_dynamic_state(runs[0], runs[1], f"{nx}x{ny} eta={eta} geq={geq} hot={hot}")
