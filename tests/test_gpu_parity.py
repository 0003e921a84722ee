"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

FP64 storage must be bitwise equal to the oracle for populations, fields,
fy_node and nu_t (same operand order, no contraction, IEEE div/sqrt on both
sides); reductions agree to 1e-12 relative (different summation order).
The reference tests of src/uniform/tests.rs are restated on the GPU path.
"""
import numpy as np
import pytest

from cfd_cavity_amd import CavitySolver
from cfd_cavity_amd._lib import CavlbmError, E_SIZE, E_UNSUPPORTED
from tests.helpers import SMALL, args_of, bits_equal, first_mismatch, rel

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "ux", "uy", "temperature", "nu_t", "fy_node")


def pair(oracle, cfg=SMALL, jobs=1, **kw):
    g = CavitySolver(*args_of(cfg), jobs, **{k: v for k, v in kw.items() if k != "oracle_jobs"})
    okw = {k: v for k, v in kw.items() if k in ("alpha_mult", "nu_start_mult", "nu_override",
                                                "geq_order")}
    o = oracle.OracleSolver(*args_of(cfg), kw.get("oracle_jobs", 1), **okw)
    return g, o


def assert_state_bitwise(g, o, what=""):
    for name in FIELDS:
        got = g.get_field(name)
        ref = getattr(o, name if name != "temperature" else "temperature")
        assert bits_equal(got, ref), f"{what} {name}: {first_mismatch(got, ref)}"
    for which in ("f", "g"):
        got = g.get_pops(which)
        ref = getattr(o, which)
        assert bits_equal(got, ref), f"{what} pops {which}: {first_mismatch(got, ref)}"


def assert_diag_close(g, o, tol=1e-12):
    dg = g.diagnostics()
    do = o.diagnostics()
    for key in ("nusselt_hot", "nusselt_cold", "nusselt_hot_wall_link", "nusselt_volume",
                "total_mass", "total_f_mass", "y_plus_hot"):
        assert rel(dg[key], do[key]) <= tol, (key, dg[key], do[key])
    for key in ("max_uy", "max_u", "nu_t_max"):  # order-independent: exact
        assert dg[key] == do[key], (key, dg[key], do[key])
    assert dg["all_finite"] == do["all_finite"]
    assert dg["time_step"] == do["time_step"]


# ---------------------------------------------------------------------------
# bitwise FP64 parity

def test_initial_state_bitwise(oracle):
    g, o = pair(oracle)
    assert_state_bitwise(g, o, "init")
    d = g.derived()
    od = o.derived()
    for k in ("lz", "lx", "nu", "alpha", "alpha_phys", "beta_eff", "tau", "tau_t", "cs", "pr_t",
              "nu_start_mult", "t_ref", "spin_up_steps"):
        assert d[k] == od[k], k


@pytest.mark.parametrize("nsteps", [1, 2, 10, 210])
def test_small_solver_bitwise(oracle, nsteps):
    g, o = pair(oracle)
    g.run(nsteps)
    o.run(nsteps)
    assert g.time_step == nsteps
    assert_state_bitwise(g, o, f"16x40 after {nsteps}")
    assert_diag_close(g, o)


def test_64x256_ra1e6_bitwise_1000_steps(oracle):
    cfg = dict(SMALL, nx=64, ny=256, ra=1.0e6)
    g, o = pair(oracle, cfg, oracle_jobs=8)
    g.run(1000)
    o.run(1000)
    assert_state_bitwise(g, o, "64x256 Ra1e6 @1000")
    assert_diag_close(g, o)


def test_128x512_ra1e8_bitwise_1k_and_10k_steps(oracle):
    """The GPU drop-in configuration: 128x512, Ra=1e8, compared at 1k and 10k
    steps (the 100k-step comparison is tests/long_parity.py)."""
    cfg = dict(SMALL, nx=128, ny=512, ra=1.0e8)
    g, o = pair(oracle, cfg, oracle_jobs=8)
    for upto in (1000, 10000):
        g.run(upto - g.time_step)
        o.run(upto - o.time_step)
        print(f"[parity] 128x512 Ra1e8 at {upto}", flush=True)
        assert_state_bitwise(g, o, f"128x512 Ra1e8 @{upto}")
        assert_diag_close(g, o)


def test_256x1024_ra1e10_bitwise_short(oracle):
    cfg = dict(SMALL, nx=256, ny=1024, ra=1.0e10, utarget=0.005)
    g, o = pair(oracle, cfg, oracle_jobs=8)
    g.run(30)
    o.run(30)
    assert_state_bitwise(g, o, "256x1024 Ra1e10 @30")
    assert_diag_close(g, o)


@pytest.mark.parametrize("kw", [dict(geq_order=1), dict(nu_start_mult=4.0),
                                dict(alpha_mult=2.5), dict(nu_override=0.01)])
def test_options_bitwise(oracle, kw):
    g, o = pair(oracle, **kw)
    g.run(230)
    o.run(230)
    assert_state_bitwise(g, o, str(kw))


def test_graph_replay_equals_single_steps(oracle):
    a = CavitySolver(*args_of(SMALL))
    b = CavitySolver(*args_of(SMALL))
    a.run(600)              # direct launches through spin-up, then graph replays
    for _ in range(600):
        b.step()
    for name in FIELDS:
        assert bits_equal(a.get_field(name), b.get_field(name)), name
    assert bits_equal(a.get_pops("f"), b.get_pops("f"))
    assert bits_equal(a.get_pops("g"), b.get_pops("g"))


_DIRECT_CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from cfd_cavity_amd import CavitySolver
from tests.helpers import SMALL, args_of
s = CavitySolver(*args_of(SMALL))
if sys.argv[3] == "dyn":
    s.set_dynamic_sgs(True)
s.run(600)
np.savez(sys.argv[2], f=s.get_pops("f"), g=s.get_pops("g"), temperature=s.get_field("temperature"),
         nu_t=s.get_field("nu_t"))
"""


@pytest.mark.parametrize("mode", ["plain", "dyn"])
def test_direct_launches_equal_graph_replay(tmp_path, mode):
    """CAVLBM_GRAPHS=0 (read once per process, so in a child process) launches
    every step block directly; the run after the spin-up, which otherwise
    replays captured graphs, gives the same bits."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = str(tmp_path / "direct.npz")
    env = dict(os.environ, CAVLBM_GRAPHS="0")
    subprocess.run([sys.executable, "-c", _DIRECT_CHILD, root, out, mode], env=env, check=True,
                   timeout=300)
    d = np.load(out)
    a = CavitySolver(*args_of(SMALL))
    if mode == "dyn":
        a.set_dynamic_sgs(True)
    assert a.spin_up_steps < 600 - 48
    a.run(600)
    for name, got in (("f", a.get_pops("f")), ("g", a.get_pops("g")),
                      ("temperature", a.get_field("temperature")), ("nu_t", a.get_field("nu_t"))):
        assert bits_equal(got, d[name]), f"{mode} {name}: {first_mismatch(got, d[name])}"


def test_collision_zero_terms_on_exact_zero_states(oracle):
    """The collision leaves out terms that are always a zero (b_g, the odd
    force part of the cy = 0 directions, the +0 starts of the sums) and takes
    (0 - uy)/cs2 as -(uy/cs2).  On states full of exact zeros -- u = 0 at
    equilibrium (zero non-equilibrium parts), uy = -0.0, T = 0 (f_y = 0),
    populations exactly 0 -- one collision must equal the oracle's: bit for
    bit where the populations are nonzero, as values everywhere."""
    g, o = pair(oracle)
    n = SMALL["nx"] * SMALL["ny"]
    rng = np.random.default_rng(11)
    rho = 1.0 + 1e-3 * rng.standard_normal(n)
    ux = 1e-3 * rng.standard_normal(n)
    uy = 1e-3 * rng.standard_normal(n)
    t = 0.5 * rng.uniform(-1, 1, n)
    kind = np.arange(n) % 5
    ux[kind == 0] = 0.0          # at rest: f = feq, g = geq, every d = 0
    uy[kind == 0] = 0.0
    uy[kind == 1] = -0.0         # negative zero velocity component
    t[kind == 2] = 0.0           # T = t_ref: f_y = 0
    ux[kind == 3] = 0.0
    uy[kind == 3] = 0.0
    t[kind == 3] = 0.0           # at rest and f_y = 0
    for s in (g, o):
        s.init_from_fields(rho, ux, uy, t)
    f = g.get_pops("f")
    gp = g.get_pops("g")
    assert bits_equal(f, o.f) and bits_equal(gp, o.g)
    # kind 4: perturbed populations, a few of them exactly zero
    f4 = f.copy()
    g4 = gp.copy()
    k4 = np.nonzero(kind == 4)[0]
    f4[k4] += 1e-4 * rng.standard_normal((k4.size, 9))
    g4[k4] += 1e-4 * rng.standard_normal((k4.size, 9))
    f4[k4[::7], 3] = 0.0
    g4[k4[::5], 6] = 0.0
    g4[k4[1::9], 0] = -0.0
    g.set_pops("f", f4)
    g.set_pops("g", g4)
    o.f[:] = f4
    o.g[:] = g4
    for name, v in (("rho", rho), ("ux", ux), ("uy", uy), ("temperature", t)):
        g.set_field(name, v)
        getattr(o, name)[:] = v
    assert bits_equal(g.get_field("uy"), o.uy)  # the -0.0 entries reached both
    g.collide(500)
    o.collide(500)
    gf, gg = g.get_pops("f"), g.get_pops("g")
    # population by population: a nonzero input keeps the reference's bits
    # (p + a zero of either sign is p); T = 0 makes every g of those nodes 0
    for got, ref, before in ((gf, o.f, f4), (gg, o.g, g4)):
        nz = before != 0.0
        assert nz.sum() > 0.5 * nz.size
        assert bits_equal(got[nz], ref[nz]), first_mismatch(got[nz], ref[nz])
        assert np.array_equal(got, ref)  # values everywhere (+0 == -0)
    for name in ("nu_t", "fy_node"):
        assert np.array_equal(g.get_field(name), getattr(o, name)), name


def test_interleaved_reads_do_not_perturb(oracle):
    g, o = pair(oracle)
    for chunk in (3, 1, 70, 200, 5):
        g.run(chunk)
        o.run(chunk)
        g.diagnostics()
        g.get_field("temperature")
    assert_state_bitwise(g, o, "interleaved")


def test_phase_split_bitwise(oracle):
    g, o = pair(oracle)
    g.run(5)
    o.run(5)
    steps = [("collide", lambda s: s.collide(100)), ("stream", lambda s: s.stream()),
             ("bounce_back", lambda s: s.apply_bounce_back()),
             ("thermal_bc", lambda s: s.apply_thermal_boundary_conditions()),
             ("update_macroscopic", lambda s: s.update_macroscopic()),
             ("enforce_wall_velocity", lambda s: s.enforce_wall_velocity())]
    for name, fn in steps:
        fn(g)
        fn(o)
        assert_state_bitwise(g, o, f"after phase {name}")
    # stepping resumes from the phase-modified state
    g.run(3)
    o.run(3)
    assert_state_bitwise(g, o, "after phases + 3 steps")


def test_init_from_fields_and_setters_bitwise(oracle):
    g, o = pair(oracle)
    n = SMALL["nx"] * SMALL["ny"]
    rng = np.random.default_rng(7)
    rho = 1.0 + 1e-3 * rng.standard_normal(n)
    ux = 1e-3 * rng.standard_normal(n)
    uy = 1e-3 * rng.standard_normal(n)
    t = 0.5 * rng.uniform(-1, 1, n)
    g.init_from_fields(rho, ux, uy, t)
    o.init_from_fields(rho, ux, uy, t)
    assert g.time_step == 0
    assert_state_bitwise(g, o, "init_from_fields")
    g.run(25)
    o.run(25)
    assert_state_bitwise(g, o, "init_from_fields + 25")
    # direct field / population writes (the reference mutates public fields)
    t2 = o.temperature.copy()
    t2[::7] += 1e-3
    g.set_field("temperature", t2)
    o.temperature[:] = t2
    f2 = o.f.copy()
    f2[:, 0] *= 1.0 + 1e-9
    g.set_pops("f", f2)
    o.f[:] = f2
    g.run(10)
    o.run(10)
    assert_state_bitwise(g, o, "after setters + 10")


def test_hot_wall_profile_bitwise(oracle):
    g, o = pair(oracle)
    rng = np.random.default_rng(11)
    theta = 0.5 + 0.01 * rng.uniform(-1, 1, SMALL["ny"])
    g.set_hot_wall(theta)
    o.set_hot_wall(theta)
    g.run(240)
    o.run(240)
    assert_state_bitwise(g, o, "hot wall profile")
    # corner rows take theta_hot[1] / theta_hot[ny-2] (solver.rs:974-987)
    tg = g.get_field("temperature").reshape(SMALL["ny"], SMALL["nx"])
    assert tg[0, 0] == tg[1, 0]


def test_errors_follow_reference_asserts():
    g = CavitySolver(*args_of(SMALL))
    g.set_dynamic_sgs(False)
    with pytest.raises(ValueError):
        g.init_from_fields(np.ones(3), np.ones(3), np.ones(3), np.ones(3))
    with pytest.raises(CavlbmError) as e:
        g._check(g._L.cavlbm_get_field(g._h, 0, None, 5))
    with pytest.raises(ValueError):
        CavitySolver(16, 40, 1e5, 0.71, 0.1, 0.9, 0.02, 42, 1, 1.0, 1.0, nu_override=-1.0)
    assert E_SIZE == -2


# ---------------------------------------------------------------------------
# reference tests (src/uniform/tests.rs) restated on the GPU path

def small():
    return CavitySolver(*args_of(SMALL))


def test_init_temperature_linear_and_bounded():   # tests.rs:138-149
    s = small()
    t = s.lbm.temperature
    assert abs(t[s.idx(0, 10)] - 0.5) < 2e-3
    assert abs(t[s.idx(s.nx() - 1, 10)] + 0.5) < 2e-3
    assert np.all(np.abs(s.lbm.ux) < 1e-10) and np.all(np.abs(s.lbm.uy) < 1e-10)


def test_mass_conserved_per_phase():               # tests.rs:154-169
    s = small()
    m0 = s.total_f_mass()
    s.collide(100)
    m1 = s.total_f_mass()
    s.stream()
    m2 = s.total_f_mass()
    s.apply_bounce_back()
    m3 = s.total_f_mass()
    s.apply_thermal_boundary_conditions()
    m4 = s.total_f_mass()
    for m in (m1, m2, m3, m4):
        assert abs(m - m0) < 1e-8


def test_guo_forcing_preserves_mass_and_upward_sign():   # tests.rs:174-222
    s = small()
    m0 = s.total_f_mass()
    s.run(210)
    m1 = s.total_f_mass()
    assert abs(m1 - m0) / abs(m0) < 1e-9
    nx, ny = s.nx(), s.ny()
    f = s.lbm.f
    rho = s.lbm.rho
    uy = s.lbm.uy
    fy = s.fy_node
    cy = np.array([0, 0, 1, 1, 1, 0, -1, -1, -1], dtype=np.float64)
    for iy in range(1, ny - 1):
        for ix in range(1, nx - 1):
            k = s.idx(ix, iy)
            mom_y = 0.0
            for j in range(9):
                mom_y += cy[j] * f[k, j]
            expected = mom_y / rho[k]
            expected += 0.5 * fy[k] / rho[k]
            assert abs(uy[k] - expected) < 1e-12


def test_hot_fluid_pushed_upward():                # tests.rs:381-401
    s = small()
    s.run(200)
    uy = s.lbm.uy
    vals = [uy[s.idx(ix, iy)] for iy in range(s.ny() // 2, s.ny() - 2) for ix in range(1, 4)]
    assert np.mean(vals) > 0.0


def test_walls_after_50_steps():                   # tests.rs:406-476
    s = small()
    s.run(50)
    nx, ny = s.nx(), s.ny()
    t = s.lbm.temperature.reshape(ny, nx)
    ux = s.lbm.ux.reshape(ny, nx)
    uy = s.lbm.uy.reshape(ny, nx)
    assert np.all(np.abs(t[:, 0] - 0.5) < 1e-9)
    assert np.all(np.abs(t[:, -1] + 0.5) < 1e-9)
    assert np.all(np.abs(t[0, :] - t[1, :]) < 1e-9)
    assert np.all(np.abs(t[-1, :] - t[-2, :]) < 1e-9)
    for a in (ux, uy):
        assert np.all(np.abs(a[:, 0]) < 1e-12) and np.all(np.abs(a[:, -1]) < 1e-12)
        assert np.all(np.abs(a[0, :]) < 1e-12) and np.all(np.abs(a[-1, :]) < 1e-12)


def test_nusselt_conduction_limit_near_az():       # tests.rs:481-506
    s = small()
    nx, ny = s.nx(), s.ny()
    t = np.array([0.5 + (-0.5 - 0.5) * (ix / (nx - 1)) for iy in range(ny) for ix in range(nx)])
    s.set_field("temperature", t)
    az = (ny - 1) / (nx - 1)
    assert abs(s.diagnostics()["nusselt_hot"] - az) < 0.05   # post.rs:68-78


def test_thermal_wall_link_conductance(oracle):    # tests.rs:602-643
    s = CavitySolver(21, 21, 1.0, 1.0, 0.0, 1.0, 0.02, 42, 1, 1.0, 1.0, nu_override=0.00169)
    s.run(20_000)
    nx, ny = s.nx(), s.ny()
    t = s.lbm.temperature.reshape(ny, nx)
    rows = slice(1, ny - 1)
    hot1 = np.mean(np.abs(t[rows, 0] - t[rows, 1]))
    hot2 = np.mean(np.abs(t[rows, 1] - t[rows, 2]))
    cold1 = np.mean(np.abs(t[rows, nx - 1] - t[rows, nx - 2]))
    cold2 = np.mean(np.abs(t[rows, nx - 2] - t[rows, nx - 3]))
    assert abs(hot1 / hot2 - 0.5) < 0.05 and abs(cold1 / cold2 - 0.5) < 0.05
    o = oracle.OracleSolver(21, 21, 1.0, 1.0, 0.0, 1.0, 0.02, 42, 1, nu_override=0.00169)
    o.run(20_000)
    assert bits_equal(s.lbm.temperature, o.temperature)


# ---------------------------------------------------------------------------
# run-loop helpers: time averaging on the device, snapshot restart

def test_time_averager_matches_host_sums(oracle):      # post.rs:11-57
    g, o = pair(oracle)
    sums = [np.zeros(SMALL["nx"] * SMALL["ny"]) for _ in range(5)]
    g.average_reset()
    for step in range(1, 301):
        g.step()
        o.step()
        if step >= 100 and step % 20 == 0:
            g.average_accumulate()
            ux, uy = o.ux.copy(), o.uy.copy()
            for a, v in zip(sums, (ux, uy, o.temperature, o.nu_t, ux * ux + uy * uy)):
                a += v
    n = g.average_count()
    assert n == 11
    inv = 1.0 / n
    for got, s in zip(g.average_means(), sums):
        assert bits_equal(got, s * inv)


def test_snapshot_restart_regrid_matches_oracle(oracle, tmp_path):   # main.rs:104-128
    from cfd_cavity_amd import snapshot

    g, o = pair(oracle)
    g.run(150)
    path = tmp_path / "snap.bin"
    snapshot.save(g, path)
    # restart on a finer grid: interpolate, init_from_fields, both backends
    big = dict(SMALL, nx=24, ny=60)
    g2, o2 = pair(oracle, big)
    snapshot.restart(g2, path)
    snx, sny, rho, ux, uy, t = snapshot.read_snapshot(path)
    assert (snx, sny) == (16, 40)
    f = [snapshot.interpolate_field(a, snx, sny, 24, 60) for a in (rho, ux, uy, t)]
    o2.init_from_fields(*f)
    assert g2.time_step == 0
    g2.run(40)
    o2.run(40)
    assert_state_bitwise(g2, o2, "after regrid restart + 40")
