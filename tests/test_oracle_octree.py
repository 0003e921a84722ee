"""Pins of the octree oracle (no GPU).

The oracle restates src/octree (cavity.rs, grid.rs, interface.rs, lbm.rs,
solver.rs).  It is pinned by
  - the reference's own octree tests (src/octree/tests.rs:135-575: the V1
    interface transfers, the V2 cavity and the V6 three-level staircase);
  - bitwise agreement of every per-level phase with the separately pinned
    uniform oracle: on the full coarse level, and on the interior of a block
    whose four sides are walls (the block rules then reduce to the uniform
    ones);
  - closed forms: bilinear stencils and both transfers are exact on linear
    equilibrium fields; the tables have the sizes the geometry fixes;
  - the initial state: the coarse temperature is the uniform solver's, and the
    fine levels hold its bilinear interpolation (an independent stencil).
"""
import math

import numpy as np
import pytest

from tests.helpers import bits_equal

D2Q9_C = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [-1, 1], [-1, 0], [-1, -1], [0, -1], [1, -1]])
CX = D2Q9_C[:, 0].astype(float)
CY = D2Q9_C[:, 1].astype(float)
W = np.array([4 / 9, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36])

# tests.rs:30-40
NX0 = NY0 = 16
BX0, BY0, BX1, BY1 = 4, 4, 12, 12
NU_COARSE, ALPHA_COARSE = 0.01, 0.02


def feq_vec(oracle, rho, ux, uy):
    return np.array([oracle.feq(i, rho, ux, uy) for i in range(9)])


def geq_vec(oracle, t, ux, uy):
    return np.array([oracle.geq(i, t, ux, uy) for i in range(9)])


# ---------------------------------------------------------------------------
# V1: interface transfers (tests.rs:63-395)

def init_v1(oracle):
    """tests.rs:63-100: Couette uy linear in x, T linear in y, rho with a sine,
    odd fneq on f and even gneq on g; moments recomputed from f."""
    s = oracle.OracleOct2(NX0, NY0, BX0, BY0, BX1, BY1, NU_COARSE, ALPHA_COARSE)
    c = s.coarse
    eps, eps_f, eps_g, u0, ct = 1e-3, 1e-4, 1e-5, 0.01, 0.01
    for iy in range(NY0):
        for ix in range(NX0):
            k = iy * NX0 + ix
            uy = u0 * (ix - NX0 / 2) / (NX0 / 2)
            temp = 0.5 + ct * (iy - NY0 / 2) / (NY0 / 2)
            rho = 1.0 + eps * math.sin(math.tau * iy / NY0)
            c.f[k] = feq_vec(oracle, rho, 0.0, uy) + eps_f * W * CY
            c.g[k] = geq_vec(oracle, temp, 0.0, uy) + eps_g * W * CX
    # LbmLevel::update_macroscopic (lbm.rs:55-73)
    c.rho[:] = c.f.sum(1)
    c.ux[:] = (c.f * CX).sum(1) / c.rho
    c.uy[:] = (c.f * CY).sum(1) / c.rho
    c.temperature[:] = c.g.sum(1)
    return s


def bilinear_at(x, y):
    ix, iy = int(math.floor(x)), int(math.floor(y))
    fx, fy = x - ix, y - iy
    return [(iy * NX0 + ix, (1 - fx) * (1 - fy)), (iy * NX0 + ix + 1, fx * (1 - fy)),
            ((iy + 1) * NX0 + ix, (1 - fx) * fy), ((iy + 1) * NX0 + ix + 1, fx * fy)]


def test_v1_coarse_to_fine_conserves_mass_and_momentum(oracle):   # tests.rs:135-200
    s = init_v1(oracle)
    s.coarse_to_fine()
    c, f = s.coarse, s.fine
    xi_odd = s.xi()[1]
    nx1 = f.nx
    errs = np.zeros(4)
    sums = np.zeros(4)
    exps = np.zeros(4)
    for k in s.ghost_ks():
        a, b = k % nx1, k // nx1
        x, y = BX0 - 0.5 + 0.5 * a, BY0 - 0.5 + 0.5 * b
        rho = ux = uy = t = fmx = fmy = 0.0
        for kc, w in bilinear_at(x, y):
            rho += w * c.rho[kc]
            ux += w * c.ux[kc]
            uy += w * c.uy[kc]
            t += w * c.temperature[kc]
            fneq = c.f[kc] - feq_vec(oracle, c.rho[kc], c.ux[kc], c.uy[kc])
            fmx += w * (CX * fneq).sum()
            fmy += w * (CY * fneq).sum()
        exp = np.array([rho, rho * ux + xi_odd * fmx, rho * uy + xi_odd * fmy, t])
        got = np.array([f.f[k].sum(), (CX * f.f[k]).sum(), (CY * f.f[k]).sum(), f.g[k].sum()])
        errs = np.maximum(errs, np.abs(got - exp))
        sums += got
        exps += exp
    assert np.all(errs < 1e-12), errs
    assert np.all(np.abs(sums - exps) < 1e-12)


def test_v1_fine_to_coarse_conserves_and_inverse(oracle):         # tests.rs:215-314
    s = oracle.OracleOct2(NX0, NY0, BX0, BY0, BX1, BY1, NU_COARSE, ALPHA_COARSE)
    f = s.fine
    eps, eps_e, eps_o, u0 = 1e-3, 2e-4, 5e-4, 0.02
    for b in range(f.ny):
        for a in range(f.nx):
            k = b * f.nx + a
            x, y = BX0 - 0.5 + 0.5 * a, BY0 - 0.5 + 0.5 * b
            uy = u0 * (x - NX0 / 2) / (NX0 / 2)
            temp = 0.5 + 0.01 * (y - NY0 / 2) / (NY0 / 2)
            f.rho[k] = 1.0 + eps * math.sin(math.tau * y / f.ny)
            f.ux[k], f.uy[k], f.temperature[k] = 0.0, uy, temp
            fneq = eps_e * W * CX * CY + eps_o * W * CX
            f.f[k] = feq_vec(oracle, f.rho[k], 0.0, uy) + fneq
            f.g[k] = geq_vec(oracle, temp, 0.0, uy) + 1e-5 * W * CY
    s.fine_to_coarse()
    xi = s.xi()
    c = s.coarse
    errs = np.zeros(4)
    for kc, kfs, ws in s.coincident():
        e = np.zeros(4)
        fmx = fmy = 0.0
        for kf, w in zip(kfs, ws):
            e += w * np.array([f.rho[kf], f.rho[kf] * f.ux[kf], f.rho[kf] * f.uy[kf],
                               f.temperature[kf]])
            fneq = f.f[kf] - feq_vec(oracle, f.rho[kf], f.ux[kf], f.uy[kf])
            fmx += w * (CX * fneq).sum()
            fmy += w * (CY * fneq).sum()
        e[1] += fmx / xi[1]
        e[2] += fmy / xi[1]
        got = np.array([c.f[kc].sum(), (CX * c.f[kc]).sum(), (CY * c.f[kc]).sum(), c.g[kc].sum()])
        errs = np.maximum(errs, np.abs(got - e))
    assert np.all(errs < 1e-12), errs
    fneq = eps_e * W * CX * CY + eps_o * W * CX
    up = oracle.oct_rescale_fneq(fneq, xi)
    down = oracle.oct_rescale_fneq(up, xi, inverse=True)
    assert np.max(np.abs(down - fneq)) < 1e-15


def test_v1_xi_matches_relaxation_time_ratio(oracle):             # tests.rs:321-395
    tc = oracle.oct_relax_times(NU_COARSE, ALPHA_COARSE)
    tf = oracle.oct_relax_times(2 * NU_COARSE, 2 * ALPHA_COARSE)
    assert abs(tf[0] - 0.5 - 2.0 * (tc[0] - 0.5)) < 1e-15
    xi = oracle.oct_rescaling_factors(tc, tf, 2, False)
    assert xi[0] == 0.5 * tf[0] / tc[0] and xi[2] == 0.5 * tf[2] / tc[2] and xi[1] == 0.5
    xe = oracle.oct_rescaling_factors(tc, tf, 2, True)
    assert xe[1] == 0.5 * tf[1] / tc[1]
    assert tc[1] == 0.5 + (3.0 / 16.0) / (tc[0] - 0.5)
    fe = 3e-3 * W * CX * CY
    fo = 5e-3 * W * CX
    assert np.max(np.abs(oracle.oct_rescale_fneq(fe, xi) - xi[0] * fe)) == 0.0
    assert np.max(np.abs(oracle.oct_rescale_fneq(fo, xi) - xi[1] * fo)) == 0.0
    assert abs(oracle.oct_rescale_fneq(fe + fo, xi).sum()) < 1e-15


# ---------------------------------------------------------------------------
# grid tables and transfers: closed forms

def test_bilinear_stencil_independent_and_clamped(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        nx0, ny0 = int(rng.integers(3, 40)), int(rng.integers(3, 40))
        x, y = rng.uniform(-0.5, nx0 - 0.5), rng.uniform(-0.5, ny0 - 0.5)
        k, w = oracle.oct_bilinear_stencil(x, y, nx0, ny0)
        ix = min(max(math.floor(x), 0), nx0 - 2)
        iy = min(max(math.floor(y), 0), ny0 - 2)
        fx, fy = x - ix, y - iy
        assert k == [iy * nx0 + ix, iy * nx0 + ix + 1, (iy + 1) * nx0 + ix, (iy + 1) * nx0 + ix + 1]
        assert np.allclose(w, [(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy],
                           rtol=0, atol=1e-15)
        # exact on a linear field, also where it extrapolates
        vals = np.array([2.0 + 0.3 * (kk % nx0) - 0.7 * (kk // nx0) for kk in k])
        assert abs((w * vals).sum() - (2.0 + 0.3 * x - 0.7 * y)) < 1e-12


def test_grid_tables_sizes_and_weights(oracle):
    o = oracle.OracleOctree(32, 64, 4, 1.0e6, levels=3)
    for gi in (1, 2, 3, 4):
        g = o.grids[gi]
        gh = o.ghosts(gi)
        assert len(gh) == 2 * (g.nx + g.ny) - 4
        assert sorted(k for k, _, _ in gh) == sorted(
            b * g.nx + a for b in range(g.ny) for a in range(g.nx)
            if a in (0, g.nx - 1) or b in (0, g.ny - 1))
        for _, _, w in gh:
            assert abs(w.sum() - 1.0) < 1e-15
        for _, kf, w in o.coincident(gi):
            assert abs(w.sum() - 1.0) < 1e-15 and len(set(kf)) == 9
    # interface A (grid.rs:169-176): coarse nodes on the domain edge are
    # excluded, rows 1..62; the hot block [0, 4) loses its wall column x = 0,
    # the cold block [27, 31) stops before the wall column x = 31
    assert len(o.coincident(1)) == 3 * 62
    assert len(o.coincident(2)) == 4 * 62
    # interface B (wall_left, wall_bottom excluded): one column of L1 nodes
    assert len(o.coincident(3)) == 1 * 125 and len(o.coincident(4)) == 125


def test_transfers_exact_on_linear_equilibrium(oracle):
    """coarse->fine and fine->coarse on equilibrium fields whose moments are
    linear in x and y: bilinear interpolation (and extrapolation) and the
    symmetric 9-node restriction reproduce the moments, so the injected
    populations are the equilibria of the exact linear moments."""
    s = oracle.OracleOct2(NX0, NY0, BX0, BY0, BX1, BY1, NU_COARSE, ALPHA_COARSE)

    def mom(x, y):
        return 1.0 + 0.002 * x - 0.001 * y, 0.003 - 0.0004 * y, -0.002 + 0.0005 * x, 0.1 * x - 0.05 * y

    c = s.coarse
    for iy in range(NY0):
        for ix in range(NX0):
            k = iy * NX0 + ix
            r, u, v, t = mom(ix, iy)
            c.rho[k], c.ux[k], c.uy[k], c.temperature[k] = r, u, v, t
            c.f[k] = feq_vec(oracle, r, u, v)
            c.g[k] = geq_vec(oracle, t, u, v)
    s.coarse_to_fine()
    f = s.fine
    for k in s.ghost_ks():
        a, b = k % f.nx, k // f.nx
        r, u, v, t = mom(BX0 - 0.5 + 0.5 * a, BY0 - 0.5 + 0.5 * b)
        assert np.max(np.abs(f.f[k] - feq_vec(oracle, r, u, v))) < 1e-14
        assert np.max(np.abs(f.g[k] - geq_vec(oracle, t, u, v))) < 1e-14
        assert abs(f.rho[k] - r) < 1e-14 and abs(f.temperature[k] - t) < 1e-14
    for b in range(f.ny):
        for a in range(f.nx):
            k = b * f.nx + a
            r, u, v, t = mom(BX0 - 0.5 + 0.5 * a, BY0 - 0.5 + 0.5 * b)
            f.rho[k], f.ux[k], f.uy[k], f.temperature[k] = r, u, v, t
            f.f[k] = feq_vec(oracle, r, u, v)
            f.g[k] = geq_vec(oracle, t, u, v)
    s.fine_to_coarse()
    for kc, _, _ in s.coincident():
        r, u, v, t = mom(kc % NX0, kc // NX0)
        assert np.max(np.abs(c.f[kc] - feq_vec(oracle, r, u, v))) < 1e-14
        assert np.max(np.abs(c.g[kc] - geq_vec(oracle, t, u, v))) < 1e-14


# ---------------------------------------------------------------------------
# per-level phases against the uniform oracle (bitwise)

def random_state(rng, lv):
    n = lv.nx * lv.ny
    lv.f[:] = W * (1.0 + 0.05 * rng.uniform(-1, 1, (n, 9)))
    lv.g[:] = W * 0.4 * rng.uniform(-1, 1, (n, 9))
    lv.rho[:] = lv.f.sum(1)
    lv.ux[:] = 0.02 * rng.uniform(-1, 1, n)
    lv.uy[:] = 0.02 * rng.uniform(-1, 1, n)
    lv.temperature[:] = lv.g.sum(1)


def uniform_of(oracle, nx, ny):
    return oracle.OracleSolver(nx, ny, 1e6, 0.71, 0.1, 0.9, 0.02, 42, 1)


def test_coarse_phases_equal_uniform_phases(oracle):
    """collide_cavity (full window, fixed taus), stream_cavity, the thermal
    BCs, update_macroscopic_guo and the wall velocities of the coarse level are
    the uniform solver's phases with Smagorinsky off: bitwise equal."""
    rng = np.random.default_rng(7)
    o = oracle.OracleOctree(32, 64, 4, 1.0e6)
    L0 = o.grids[0]
    random_state(rng, L0)
    u = uniform_of(oracle, 32, 64)
    for name in ("f", "g", "rho", "ux", "uy", "temperature"):
        getattr(u, name)[:] = getattr(L0, name)
    d = o.derived()
    tau = [d["tau_plus_l0"], d["tau_minus_l0"], d["tau_t_l0"]]
    beta = d["beta_eff_coarse"]
    o.collide_level(0, tau, beta, False)
    for k in range(32 * 64):
        u.f[k], u.g[k], _, u.fy_node[k] = oracle.collide_node(
            u.f[k], u.g[k], u.rho[k], u.ux[k], u.uy[k], u.temperature[k], 0.5, beta, 0.0, 0.0,
            1.0, d["nu_lu"], d["alpha_lu"])
    assert bits_equal(L0.f, u.f) and bits_equal(L0.g, u.g) and bits_equal(L0.fy, u.fy_node)
    o.stream_level(0)
    u.stream()
    assert bits_equal(L0.f, u.f) and bits_equal(L0.g, u.g)
    o.bc_level(0)
    u.apply_thermal_boundary_conditions()
    assert bits_equal(L0.g, u.g)
    o.macro_level(0)
    u.update_macroscopic()
    o.walls_level(0)
    u.enforce_wall_velocity()
    for name in ("rho", "ux", "uy", "temperature"):
        assert bits_equal(getattr(L0, name), getattr(u, name)), name


@pytest.mark.parametrize("gi", [1, 3])
def test_walled_block_interior_equals_uniform(oracle, gi):
    """A block level whose four sides are walls (stream flags all set, both
    isothermal columns) evolves its interior [1, n-2]^2 exactly as a uniform
    cavity of that size: collision on the window, streaming with bounce-back
    from the wall-side ghost layer, the block BCs and wall velocities."""
    rng = np.random.default_rng(11 + gi)
    o = oracle.OracleOctree(32, 64, 4, 1.0e6, levels=3)
    B = o.grids[gi]
    random_state(rng, B)
    nx, ny = B.nx - 2, B.ny - 2
    u = uniform_of(oracle, nx, ny)

    def interior(a):
        return a.reshape(B.ny, B.nx, *a.shape[1:])[1:-1, 1:-1].reshape(nx * ny, *a.shape[1:])

    for name in ("f", "g", "rho", "ux", "uy", "temperature"):
        getattr(u, name)[:] = interior(getattr(B, name))
    B.fy[:] = 0.0
    ghosts_before = [B.f.copy(), B.g.copy()]
    tau = oracle.oct_relax_times(0.03, 0.04)
    beta = 3e-4
    o.collide_level(gi, tau, beta, True)
    for k in range(nx * ny):
        u.f[k], u.g[k], _, u.fy_node[k] = oracle.collide_node(
            u.f[k], u.g[k], u.rho[k], u.ux[k], u.uy[k], u.temperature[k], 0.5, beta, 0.0, 0.0,
            1.0, 0.03, 0.04)
    assert bits_equal(interior(B.f), u.f) and bits_equal(interior(B.g), u.g)
    edge = np.ones((B.ny, B.nx), bool)
    edge[1:-1, 1:-1] = False
    edge = edge.ravel()
    assert bits_equal(B.f[edge], ghosts_before[0][edge])  # ghosts are not collided
    o.stream_block_level(gi, True, True, True, True)
    u.stream()
    assert bits_equal(interior(B.f), u.f) and bits_equal(interior(B.g), u.g)
    o.bc_block_level(gi, True, True)
    u.apply_thermal_boundary_conditions()
    assert bits_equal(interior(B.g), u.g)
    o.macro_level(gi)
    u.fy_node[:] = interior(B.fy)
    u.update_macroscopic()
    o.walls_block_level(gi, True, True)
    u.enforce_wall_velocity()
    for name in ("rho", "ux", "uy", "temperature"):
        assert bits_equal(interior(getattr(B, name)), getattr(u, name)), name


def test_block_bc_ghost_layers_copy_inwards(oracle):
    """cavity.rs:584-600: after the block BCs the ghost columns and rows hold
    the adjacent inner layer (zero gradient), corners included."""
    rng = np.random.default_rng(3)
    o = oracle.OracleOctree(32, 64, 4, 1.0e6)
    B = o.grids[1]
    random_state(rng, B)
    o.bc_block_level(1, True, False)
    g = B.g.reshape(B.ny, B.nx, 9)
    assert bits_equal(g[:, 0], g[:, 1]) and bits_equal(g[:, -1], g[:, -2])
    assert bits_equal(g[0], g[1]) and bits_equal(g[-1], g[-2])
    assert np.allclose(g[5, 1], W * 0.5, rtol=0, atol=1e-16)  # hot wall column a = 1


# ---------------------------------------------------------------------------
# the solver: the reference's V2 / V6 tests and the initial state

def test_v2_mass_conserved_with_guo_force(oracle):                 # tests.rs:396-416
    o = oracle.OracleOctree(32, 64, 4, 1.0e6)
    for _ in range(10):
        o.step()
        assert abs(o.grids[0].rho.sum() / (32 * 64) - 1.0) < 1e-6


def test_v2_noise_stats_bounded(oracle):                           # tests.rs:418-447
    o = oracle.OracleOctree(32, 64, 4, 1.0e6)
    for step in range(20):
        o.step()
        n = o.noise_stats()
        assert n["max_rho_dev"] >= n["max_rho_dev_ifaces"]
        if step >= 5:
            assert n["max_rho_dev_ifaces"] < 1.0


def test_v2_nusselt_conductive_initial(oracle):                    # tests.rs:449-475
    o = oracle.OracleOctree(32, 64, 4, 1.0e6)
    assert abs(o.nusselt_hot() - 1.0) < 0.05 and abs(o.nusselt_cold() - 1.0) < 0.05


def test_v6_multilevel_mapping_structure(oracle):                  # tests.rs:477-546
    o = oracle.OracleOctree(32, 64, 4, 1.0e6, levels=3)
    d = o.derived()
    nu0 = d["nu_lu"]
    t2 = oracle.oct_relax_times(4 * nu0, 4 * nu0 / 0.71)
    assert abs(d["tau_plus_l2"] - t2[0]) < 1e-15
    assert abs(d["xi_even_b"] - 0.5 * t2[0] / d["tau_plus_l1"]) < 1e-15
    assert abs(d["tau_plus_l1"] - (0.5 + 3.0 * 2.0 * nu0)) < 1e-15
    for gi, hot in ((3, True), (4, False)):
        assert o.grids[gi].nx == 7
        parent = o.grids[gi - 2]
        coinc = o.coincident(gi)
        assert coinc
        sbx0 = 1 if hot else parent.nx - 4
        bx0_block = 0 if hot else 32 - 1 - 4
        for kc, _, _ in coinc:
            cx, cy = kc % parent.nx, kc // parent.nx
            assert cx != sbx0 and cy != 1
            x_phys = bx0_block - 0.5 + 0.5 * cx
            dist = x_phys if hot else (32 - 1) - x_phys
            assert dist <= 1.0 + 1e-12


def test_v6_three_levels_stable_short(oracle):                     # tests.rs:548-575
    o = oracle.OracleOctree(32, 64, 4, 1.0e6, levels=3)
    for _ in range(10):
        o.step()
        n = o.noise_stats()
        assert abs(o.grids[0].rho.sum() / (32 * 64) - 1.0) < 1e-6
        assert n["max_rho_dev"] >= n["max_rho_dev_ifaces"] and n["max_rho_dev_ifaces"] < 1.0
        assert o.check_guards() == 0


def test_initial_state(oracle):
    """init_cavity (solver.rs:500-575): the coarse temperature is the uniform
    solver's initial field (same noise sequence), the fine levels hold its
    bilinear interpolation at their node positions, populations are equilibria."""
    o = oracle.OracleOctree(32, 64, 4, 1.0e6, levels=3)
    u = oracle.OracleSolver(32, 64, 1e6, 0.71, 0.1, 0.9, 0.02, 42, 1)
    assert bits_equal(o.grids[0].temperature, u.temperature)
    tc = o.grids[0].temperature

    def interp(x, y):
        ix = min(max(math.floor(x), 0), 30)
        iy = min(max(math.floor(y), 0), 62)
        fx, fy = x - ix, y - iy
        v = tc.reshape(64, 32)
        return ((1 - fx) * (1 - fy) * v[iy, ix] + fx * (1 - fy) * v[iy, ix + 1]
                + (1 - fx) * fy * v[iy + 1, ix] + fx * fy * v[iy + 1, ix + 1])

    for gi, bx0 in ((1, 0), (2, 27)):
        g = o.grids[gi]
        for b in range(0, g.ny, 7):
            for a in range(g.nx):
                k = b * g.nx + a
                assert abs(g.temperature[k] - interp(bx0 - 0.5 + 0.5 * a, -0.5 + 0.5 * b)) < 1e-15
                assert np.max(np.abs(g.g[k] - geq_vec(oracle, g.temperature[k], 0.0, 0.0))) < 1e-17
    s = o.grids[3]
    k = 3 * s.nx + 2
    x = 0 - 0.5 + 0.5 * (1 - 0.5 + 0.5 * 2)
    y = 0 - 0.5 + 0.5 * (1 - 0.5 + 0.5 * 3)
    assert abs(s.temperature[k] - interp(x, y)) < 1e-15


def test_set_from_snapshot_semantics(oracle):
    """set_from_snapshot (solver.rs:1095-1175): equilibria of the given fields,
    then u = 0 on the walls of every level, and time_step = 0."""
    o = oracle.OracleOctree(32, 64, 4, 1.0e6)
    o.run(3)
    rng = np.random.default_rng(9)
    sizes = [g.nx * g.ny for g in o.grids]
    tot = sum(sizes)
    rho = 1 + 0.01 * rng.uniform(-1, 1, tot)
    ux, uy = 0.01 * rng.uniform(-1, 1, tot), 0.01 * rng.uniform(-1, 1, tot)
    t = rng.uniform(-0.5, 0.5, tot)
    o.set_from_snapshot(rho, ux, uy, t)
    assert o.time_step == 0
    off = 0
    for gi, g in enumerate(o.grids):
        n = sizes[gi]
        sl = slice(off, off + n)
        assert bits_equal(g.rho, rho[sl]) and bits_equal(g.temperature, t[sl])
        k = n // 2 + 3
        assert bits_equal(g.f[k], feq_vec(oracle, rho[sl][k], ux[sl][k], uy[sl][k]))
        U = g.ux.reshape(g.ny, g.nx)
        if gi == 0:
            assert not U[0].any() and not U[:, 0].any() and not U[:, -1].any()
        else:
            assert not U[1].any() and not U[-2].any()
            assert not U[:, 1 if gi == 1 else -2].any()
        off += n
    with pytest.raises(ValueError):
        o.set_from_snapshot(rho[:-1], ux[:-1], uy[:-1], t[:-1])


def test_construction_asserts(oracle):                              # solver.rs:356-361
    with pytest.raises(ValueError):
        oracle.OracleOctree(32, 64, 1, 1.0e6)
    with pytest.raises(ValueError):
        oracle.OracleOctree(16, 64, 8, 1.0e6)
    with pytest.raises(ValueError):
        oracle.OracleOctree(32, 64, 4, 1.0e6, levels=4)
