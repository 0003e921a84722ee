"""ctypes binding of the CPU oracle (oracle/libcavity_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg, as the checker or the timed CPU baseline.  The
product package (``cfd_cavity_amd``) never imports it.

The class mirrors the reference's ``CavitySolver`` (src/uniform/solver.rs:35-104,
:108-293) with numpy views onto the oracle's own buffers: ``f``/``g`` are
``(ny*nx, 9)`` AoS arrays, the fields are ``(ny*nx,)``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcavity_oracle.so")
_lib = None

RHO, UX, UY, T, NU_T, FY = range(6)


def build() -> None:
    """Compile the oracle with its Makefile (gcc -O3 -ffp-contract=off)."""
    import subprocess

    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
    L = C.CDLL(_LIB_PATH)
    vp, sz, u64, d, i = C.c_void_p, C.c_size_t, C.c_uint64, C.c_double, C.c_int
    dp = C.POINTER(C.c_double)
    sig = {
        "orc_new": (vp, [sz, sz, d, d, d, d, d, u64, sz, d, d, i, d, i, C.c_char_p, sz]),
        "orc_free": (None, [vp]),
        "orc_set_jobs": (None, [vp, sz]),
        "orc_step": (None, [vp]),
        "orc_run": (None, [vp, u64]),
        "orc_collide": (None, [vp, u64]),
        "orc_stream": (None, [vp]),
        "orc_apply_bounce_back": (None, [vp]),
        "orc_thermal_bc": (None, [vp]),
        "orc_update_macroscopic": (None, [vp]),
        "orc_enforce_wall_velocity": (None, [vp]),
        "orc_init_from_fields": (i, [vp, dp, dp, dp, dp, sz]),
        "orc_set_hot_wall": (None, [vp, dp]),
        "orc_set_cs": (None, [vp, d]),
        "orc_set_dynamic_sgs": (None, [vp, i, d, d, d]),
        # octree (octree_oracle.h)
        "orc_oct_new": (vp, [sz, sz, sz, d, d, d, u64, sz, i, i, d, i, d, i, C.c_char_p, sz]),
        "orc_oct_free": (None, [vp]),
        "orc_oct_step": (None, [vp]),
        "orc_oct_run": (None, [vp, u64]),
        "orc_oct_ngrids": (sz, [vp]),
        "orc_oct_grid_dims": (None, [vp, sz, C.POINTER(sz), C.POINTER(sz)]),
        "orc_oct_field": (dp, [vp, sz, i]),
        "orc_oct_pops": (dp, [vp, sz, i]),
        "orc_oct_time_step": (u64, [vp]),
        "orc_oct_derived": (None, [vp, dp]),
        "orc_oct_noise_stats": (None, [vp, dp]),
        "orc_oct_max_u_all": (d, [vp]),
        "orc_oct_check_guards": (i, [vp]),
        "orc_oct_nusselt_hot": (d, [vp]),
        "orc_oct_nusselt_cold": (d, [vp]),
        "orc_oct_nusselt_hot_link0": (d, [vp]),
        "orc_oct_nusselt_hot_fine": (d, [vp, sz, d]),
        "orc_oct_set_from_snapshot": (i, [vp, dp, dp, dp, dp, sz]),
        "orc_oct_n_ghosts": (sz, [vp, sz]),
        "orc_oct_n_coincident": (sz, [vp, sz]),
        "orc_oct_ghost": (None, [vp, sz, sz, C.POINTER(sz), C.POINTER(sz), dp]),
        "orc_oct_coincident": (None, [vp, sz, sz, C.POINTER(sz), C.POINTER(sz), dp]),
        "orc_oct_collide_level": (None, [vp, sz, dp, d, i]),
        "orc_oct_stream_level": (None, [vp, sz]),
        "orc_oct_stream_block_level": (None, [vp, sz, i, i, i, i]),
        "orc_oct_bc_level": (None, [vp, sz]),
        "orc_oct_bc_block_level": (None, [vp, sz, i, i]),
        "orc_oct_macro_level": (None, [vp, sz]),
        "orc_oct_walls_level": (None, [vp, sz]),
        "orc_oct_walls_block_level": (None, [vp, sz, i, i]),
        "orc_oct_set_dynamic_sgs": (None, [vp, i, d, d, d]),
        "orc_oct_dynamic_stats": (None, [vp, dp]),
        "orc_oct_sgs_pre": (dp, [vp, sz, i]),
        "orc_oct_xi": (None, [vp, dp]),
        "orc_oct_collide_node": (None, [dp, dp, dp, d, d, d, d, dp, d, i, i, d, d]),
        "orc_dynamic_model": (None, [sz, sz, dp, dp, dp, dp, dp, d, d, d, d, d, d, dp, dp, dp]),
        "orc_oct2_new": (vp, [sz, sz, sz, sz, sz, sz, d, d, i]),
        "orc_oct2_free": (None, [vp]),
        "orc_oct2_dims": (None, [vp, C.POINTER(sz), C.POINTER(sz)]),
        "orc_oct2_field": (dp, [vp, i, i]),
        "orc_oct2_pops": (dp, [vp, i, i]),
        "orc_oct2_xi": (None, [vp, dp]),
        "orc_oct2_coarse_to_fine": (None, [vp]),
        "orc_oct2_fine_to_coarse": (None, [vp]),
        "orc_oct2_n_ghosts": (sz, [vp]),
        "orc_oct2_n_coincident": (sz, [vp]),
        "orc_oct2_ghost_k": (sz, [vp, sz]),
        "orc_oct2_coincident": (None, [vp, sz, C.POINTER(sz), C.POINTER(sz), dp]),
        "orc_oct_relax_times": (None, [d, d, dp]),
        "orc_oct_rescaling_factors": (None, [dp, dp, sz, i, dp]),
        "orc_oct_rescale_fneq": (None, [dp, dp, dp]),
        "orc_oct_rescale_fneq_inverse": (None, [dp, dp, dp]),
        "orc_oct_bilinear_stencil": (None, [d, d, sz, sz, C.POINTER(sz), dp]),
        "orc_compute_dynamic_sgs": (None, [vp]),
        "orc_lilly_coefficient": (d, [d, d]),
        "orc_mixed_factor": (d, [d, d, d]),
        "orc_dynamic_stats": (None, [vp, dp]),
        "orc_sgs_pre": (dp, [vp, i]),
        "orc_field": (dp, [vp, i]),
        "orc_pops": (dp, [vp, i]),
        "orc_nx": (sz, [vp]),
        "orc_ny": (sz, [vp]),
        "orc_time_step": (u64, [vp]),
        "orc_set_time_step": (None, [vp, u64]),
        "orc_derived": (None, [vp, dp]),
        "orc_step_scalars": (None, [vp, u64, dp]),
        "orc_total_mass": (d, [vp]),
        "orc_total_f_mass": (d, [vp]),
        "orc_nusselt_hot": (d, [vp]),
        "orc_nusselt_hot_wall_link": (d, [vp]),
        "orc_nusselt_cold": (d, [vp]),
        "orc_nusselt_volume": (d, [vp]),
        "orc_max_uy": (d, [vp]),
        "orc_max_uy_on_row": (d, [vp, sz, dp]),
        "orc_max_u": (d, [vp]),
        "orc_nu_t_max": (d, [vp]),
        "orc_wall_shear_hot": (d, [vp, sz]),
        "orc_y_plus_hot": (d, [vp]),
        "orc_check_finite": (i, [vp]),
        "orc_feq": (d, [i, d, d, d]),
        "orc_geq": (d, [i, d, d, d, i]),
        "orc_lattice": (None, [C.POINTER(C.c_int)] * 3 + [dp]),
        "orc_collide_node": (None, [dp, dp, dp, dp] + [d] * 11 + [i]),
        "orc_poiseuille_channel": (d, [sz, sz, d, d, u64]),
        "orc_rng_next_u64": (u64, [C.POINTER(C.c_uint64)]),
        "orc_rng_next_f64": (d, [C.POINTER(C.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def lilly_coefficient(num, den):
    return lib().orc_lilly_coefficient(num, den)


def mixed_factor(a, b, alpha):
    return lib().orc_mixed_factor(a, b, alpha)


def lattice():
    L = lib()
    cx = (C.c_int * 9)()
    cy = (C.c_int * 9)()
    opp = (C.c_int * 9)()
    w = np.zeros(9)
    L.orc_lattice(cx, cy, opp, _dp(w))
    return np.array(cx[:]), np.array(cy[:]), np.array(opp[:]), w


def feq(i, rho, ux, uy):
    return lib().orc_feq(i, rho, ux, uy)


def geq(i, t, ux, uy, order=2):
    return lib().orc_geq(i, t, ux, uy, order)


class OracleSolver:
    """The reference ``CavitySolver`` restated on the CPU (see module doc)."""

    def __init__(self, nx, ny, ra, pr, cs, pr_t, utarget, seed, jobs=1,
                 alpha_mult=1.0, nu_start_mult=1.0, nu_override=None, geq_order=2):
        L = lib()
        err = C.create_string_buffer(512)
        has = 0 if nu_override is None else 1
        h = L.orc_new(nx, ny, ra, pr, cs, pr_t, utarget, seed, jobs, alpha_mult,
                      nu_start_mult, has, 0.0 if nu_override is None else nu_override,
                      geq_order, err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self._h = h
        self._L = L
        self.nx, self.ny = nx, ny
        n = nx * ny
        self._n = n

        def view(ptr, shape):
            return np.ctypeslib.as_array(ptr, shape=shape)

        self.rho = view(L.orc_field(h, RHO), (n,))
        self.ux = view(L.orc_field(h, UX), (n,))
        self.uy = view(L.orc_field(h, UY), (n,))
        self.temperature = view(L.orc_field(h, T), (n,))
        self.nu_t = view(L.orc_field(h, NU_T), (n,))
        self.fy_node = view(L.orc_field(h, FY), (n,))
        self.f = view(L.orc_pops(h, 0), (n, 9))
        self.g = view(L.orc_pops(h, 1), (n, 9))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_free(h)
            self._h = None

    # reference API (src/uniform/solver.rs)
    def idx(self, ix, iy):
        return iy * self.nx + ix

    def step(self):
        self._L.orc_step(self._h)

    def run(self, n):
        self._L.orc_run(self._h, n)

    def set_jobs(self, jobs):
        self._L.orc_set_jobs(self._h, jobs)

    def collide(self, step):
        self._L.orc_collide(self._h, step)

    def stream(self):
        self._L.orc_stream(self._h)

    def apply_bounce_back(self):
        self._L.orc_apply_bounce_back(self._h)

    def apply_thermal_boundary_conditions(self):
        self._L.orc_thermal_bc(self._h)

    def update_macroscopic(self):
        self._L.orc_update_macroscopic(self._h)

    def enforce_wall_velocity(self):
        self._L.orc_enforce_wall_velocity(self._h)

    def init_from_fields(self, rho, ux, uy, t):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, ux, uy, t)]
        if any(a.size != self._n for a in arrs):
            raise ValueError("init_from_fields: field size mismatch")
        self._L.orc_init_from_fields(self._h, *[_dp(a) for a in arrs], self._n)

    def set_hot_wall(self, theta):
        if theta is None:
            self._L.orc_set_hot_wall(self._h, None)
        else:
            a = np.ascontiguousarray(theta, dtype=np.float64)
            assert a.size == self.ny
            self._L.orc_set_hot_wall(self._h, _dp(a))

    def set_cs(self, cs):
        self._L.orc_set_cs(self._h, float(cs))

    def set_dynamic_sgs(self, on, smag_alpha=0.5, nu_sm_min=-0.1, eta_test=2.0):
        self._L.orc_set_dynamic_sgs(self._h, 1 if on else 0, smag_alpha, nu_sm_min, eta_test)

    def compute_dynamic_sgs(self):
        self._L.orc_compute_dynamic_sgs(self._h)

    def dynamic_stats(self):
        out = np.zeros(4)
        self._L.orc_dynamic_stats(self._h, _dp(out))
        return tuple(float(v) for v in out)

    def sgs_pre(self, which):
        """Copy of nu_t_pre (0) / alpha_t_pre (1), or None before the model ran."""
        p = self._L.orc_sgs_pre(self._h, which)
        if not p:
            return None
        return np.ctypeslib.as_array(p, shape=(self._n,)).copy()

    @property
    def time_step(self):
        return self._L.orc_time_step(self._h)

    @time_step.setter
    def time_step(self, v):
        self._L.orc_set_time_step(self._h, v)

    def derived(self):
        out = np.zeros(13)
        self._L.orc_derived(self._h, _dp(out))
        keys = ["lz", "lx", "nu", "alpha", "alpha_phys", "beta_eff", "tau", "tau_t",
                "spin_up_steps", "cs", "pr_t", "nu_start_mult", "t_ref"]
        return dict(zip(keys, out.tolist()))

    def step_scalars(self, step):
        out = np.zeros(4)
        self._L.orc_step_scalars(self._h, step, _dp(out))
        return dict(zip(["beta", "nu_eff", "alpha_eff", "tau"], out.tolist()))

    def total_mass(self):
        return self._L.orc_total_mass(self._h)

    def total_f_mass(self):
        return self._L.orc_total_f_mass(self._h)

    def diagnostics(self):
        L, h = self._L, self._h
        return {
            "nusselt_hot": L.orc_nusselt_hot(h),
            "nusselt_cold": L.orc_nusselt_cold(h),
            "nusselt_hot_wall_link": L.orc_nusselt_hot_wall_link(h),
            "nusselt_volume": L.orc_nusselt_volume(h),
            "total_mass": L.orc_total_mass(h),
            "total_f_mass": L.orc_total_f_mass(h),
            "max_uy": L.orc_max_uy(h),
            "max_u": L.orc_max_u(h),
            "nu_t_max": L.orc_nu_t_max(h),
            "y_plus_hot": L.orc_y_plus_hot(h),
            "all_finite": bool(L.orc_check_finite(h)),
            "time_step": L.orc_time_step(h),
        }

    def max_uy_on_row(self, iy):
        x = C.c_double(0.0)
        v = self._L.orc_max_uy_on_row(self._h, iy, C.byref(x))
        return v, x.value

    def wall_shear_hot(self, iy):
        return self._L.orc_wall_shear_hot(self._h, iy)


def collide_node(f, g, rho, ux, uy, temp, tau_mol, beta, t_ref, cs, pr_t, nu, alpha,
                 geq_order=2):
    f = np.ascontiguousarray(f, dtype=np.float64).copy()
    g = np.ascontiguousarray(g, dtype=np.float64).copy()
    nut = np.zeros(1)
    fy = np.zeros(1)
    lib().orc_collide_node(_dp(f), _dp(g), _dp(nut), _dp(fy), rho, ux, uy, temp, tau_mol,
                           beta, t_ref, cs, pr_t, nu, alpha, geq_order)
    return f, g, nut[0], fy[0]


def poiseuille_channel(nx, ny, nu, fy, nsteps):
    return lib().orc_poiseuille_channel(nx, ny, nu, fy, nsteps)


# ---------------------------------------------------------------------------
# octree oracle (octree_oracle.c): the reference's src/octree restated

OCT_DERIVED = ("nu_lu", "alpha_lu", "beta_eff_coarse", "beta_eff_fine", "spin_up_steps",
               "tau_plus_l0", "tau_minus_l0", "tau_t_l0", "tau_plus_l1", "tau_minus_l1",
               "tau_t_l1", "tau_plus_l2", "tau_minus_l2", "tau_t_l2", "xi_even_a", "xi_odd_a",
               "xi_thermal_a", "xi_even_b", "xi_odd_b", "xi_thermal_b", "n_nodes_total",
               "levels")


def _view(ptr, shape):
    return np.ctypeslib.as_array(ptr, shape=shape)


def dynamic_model(f9, rho, ux, uy, temp, tau, smag_alpha, eta_test, nu_mol, alpha_mol,
                  nu_sm_min):
    """The oracle's model on given arrays (ny x nx fields, f as [k][9]):
    (nu_t, alpha_t, (cs_mean, cs_max, kappa_mean, kappa_max))."""
    ny, nx = rho.shape
    arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (f9, rho, ux, uy, temp)]
    nu_t, al_t, st = np.zeros(nx * ny), np.zeros(nx * ny), np.zeros(4)
    lib().orc_dynamic_model(nx, ny, *[_dp(a) for a in arr], tau, smag_alpha, eta_test, nu_mol,
                            alpha_mol, nu_sm_min, _dp(nu_t), _dp(al_t), _dp(st))
    return nu_t.reshape(ny, nx), al_t.reshape(ny, nx), tuple(float(v) for v in st)


def oct_collide_node(f, g, rho, ux, uy, temp, tau3, beta, geq_order=2, pre=None):
    """collide_node_cavity on copies of f, g; pre = (nu_t, alpha_t) or None."""
    f, g = np.array(f, dtype=np.float64), np.array(g, dtype=np.float64)
    fy = np.zeros(1)
    t = np.array(tau3, dtype=np.float64)
    nt, at = pre if pre is not None else (0.0, 0.0)
    lib().orc_oct_collide_node(_dp(f), _dp(g), _dp(fy), rho, ux, uy, temp, _dp(t), beta,
                               geq_order, 0 if pre is None else 1, nt, at)
    return f, g, float(fy[0])


class OracleLevel:
    """numpy views of one level: fields [n], pops [n, 9] (k = iy*nx + ix)."""

    def __init__(self, L, h, gi, field_fn, pops_fn, nx, ny):
        n = nx * ny
        self.nx, self.ny = nx, ny
        self.rho = _view(field_fn(h, gi, 0), (n,))
        self.ux = _view(field_fn(h, gi, 1), (n,))
        self.uy = _view(field_fn(h, gi, 2), (n,))
        self.temperature = _view(field_fn(h, gi, 3), (n,))
        self.f = _view(pops_fn(h, gi, 0), (n, 9))
        self.g = _view(pops_fn(h, gi, 1), (n, 9))


class OracleOctree:
    """The reference's octree ``CavitySolver`` (src/octree/solver.rs) on the
    CPU.  ``grids[gi]`` in snapshot order: 0 coarse, 1 hot L1, 2 cold L1,
    3 hot L2, 4 cold L2."""

    def __init__(self, nx0, ny0, bl_width, ra, pr=0.71, utarget=0.02, seed=42, jobs=1,
                 odd_exact=False, levels=2, nu_start_mult=1.0, nu_override=None, geq_order=2):
        L = lib()
        err = C.create_string_buffer(512)
        h = L.orc_oct_new(nx0, ny0, bl_width, ra, pr, utarget, seed, jobs, int(odd_exact),
                          levels, nu_start_mult, 0 if nu_override is None else 1,
                          0.0 if nu_override is None else nu_override, geq_order, err, 512)
        if not h:
            raise ValueError(err.value.decode())
        self._h, self._L = h, L
        self.nx0, self.ny0 = nx0, ny0
        self.grids = []
        for gi in range(L.orc_oct_ngrids(h)):
            nx, ny = C.c_size_t(), C.c_size_t()
            L.orc_oct_grid_dims(h, gi, C.byref(nx), C.byref(ny))
            lv = OracleLevel(L, h, gi, L.orc_oct_field, L.orc_oct_pops, nx.value, ny.value)
            lv.fy = _view(L.orc_oct_field(h, gi, 4), (nx.value * ny.value,))
            self.grids.append(lv)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_oct_free(h)
            self._h = None

    # -- dynamic mixed-scale SGS (solver.rs:476-481, :647-947) ---------------
    def set_dynamic_sgs(self, on=True, smag_alpha=0.5, nu_sm_min=-0.1, eta_test=2.0):
        self._L.orc_oct_set_dynamic_sgs(self._h, int(on), smag_alpha, nu_sm_min, eta_test)

    def dynamic_stats(self):
        """(cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max, nu_t_max)."""
        out = np.zeros(5)
        self._L.orc_oct_dynamic_stats(self._h, _dp(out))
        return tuple(float(v) for v in out)

    def sgs_pre(self, gi, which=0):
        """Copy of grid gi's nu_t_pre (0) or alpha_t_pre (1)."""
        g = self.grids[gi]
        return _view(self._L.orc_oct_sgs_pre(self._h, gi, which), (g.nx * g.ny,)).copy()

    def xi(self):
        """Interface A (even, odd, thermal), then B of the hot and cold block."""
        out = np.zeros(9)
        self._L.orc_oct_xi(self._h, _dp(out))
        return out

    def step(self):
        self._L.orc_oct_step(self._h)

    def run(self, n):
        self._L.orc_oct_run(self._h, n)

    @property
    def time_step(self):
        return self._L.orc_oct_time_step(self._h)

    def derived(self):
        out = np.zeros(22)
        self._L.orc_oct_derived(self._h, _dp(out))
        return dict(zip(OCT_DERIVED, out.tolist()))

    def noise_stats(self):
        out = np.zeros(3)
        self._L.orc_oct_noise_stats(self._h, _dp(out))
        return dict(max_rho_dev=out[0], max_rho_dev_ifaces=out[1], masse=out[2])

    def max_u_all(self):
        return self._L.orc_oct_max_u_all(self._h)

    def check_guards(self):
        return self._L.orc_oct_check_guards(self._h)

    def nusselt_hot(self):
        return self._L.orc_oct_nusselt_hot(self._h)

    def nusselt_cold(self):
        return self._L.orc_oct_nusselt_cold(self._h)

    def nusselt_hot_link0(self):
        return self._L.orc_oct_nusselt_hot_link0(self._h)

    def nusselt_hot_fine(self, block, lx_coarse):
        return self._L.orc_oct_nusselt_hot_fine(self._h, block, lx_coarse)

    def set_from_snapshot(self, rho, ux, uy, t):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, ux, uy, t)]
        rc = self._L.orc_oct_set_from_snapshot(self._h, *[_dp(a) for a in arrs], arrs[0].size)
        if rc:
            raise ValueError("set_from_snapshot: size mismatch")

    def ghosts(self, gi):
        out = []
        for j in range(self._L.orc_oct_n_ghosts(self._h, gi)):
            kf = C.c_size_t()
            kc = (C.c_size_t * 4)()
            w = np.zeros(4)
            self._L.orc_oct_ghost(self._h, gi, j, C.byref(kf), kc, _dp(w))
            out.append((kf.value, list(kc), w))
        return out

    def coincident(self, gi):
        out = []
        for j in range(self._L.orc_oct_n_coincident(self._h, gi)):
            k = C.c_size_t()
            kf = (C.c_size_t * 9)()
            w = np.zeros(9)
            self._L.orc_oct_coincident(self._h, gi, j, C.byref(k), kf, _dp(w))
            out.append((k.value, list(kf), w))
        return out

    # per-level phases (cavity.rs)
    def collide_level(self, gi, tau, beta, interior_only):
        t = np.ascontiguousarray(tau, dtype=np.float64)
        self._L.orc_oct_collide_level(self._h, gi, _dp(t), beta, int(interior_only))

    def stream_level(self, gi):
        self._L.orc_oct_stream_level(self._h, gi)

    def stream_block_level(self, gi, wl, wr, wb, wt):
        self._L.orc_oct_stream_block_level(self._h, gi, int(wl), int(wr), int(wb), int(wt))

    def bc_level(self, gi):
        self._L.orc_oct_bc_level(self._h, gi)

    def bc_block_level(self, gi, hot, cold):
        self._L.orc_oct_bc_block_level(self._h, gi, int(hot), int(cold))

    def macro_level(self, gi):
        self._L.orc_oct_macro_level(self._h, gi)

    def walls_level(self, gi):
        self._L.orc_oct_walls_level(self._h, gi)

    def walls_block_level(self, gi, hot, cold):
        self._L.orc_oct_walls_block_level(self._h, gi, int(hot), int(cold))


class OracleOct2:
    """The two-level interface harness ``OctreeSolver`` (solver.rs:66-116)."""

    def __init__(self, nx0, ny0, bx0, by0, bx1, by1, nu, alpha, odd_exact=False):
        L = lib()
        h = L.orc_oct2_new(nx0, ny0, bx0, by0, bx1, by1, nu, alpha, int(odd_exact))
        if not h:
            raise ValueError("bloc fin hors domaine grossier / vide")
        self._h, self._L = h, L
        nx1, ny1 = C.c_size_t(), C.c_size_t()
        L.orc_oct2_dims(h, C.byref(nx1), C.byref(ny1))
        self.coarse = OracleLevel(L, h, 0, L.orc_oct2_field, L.orc_oct2_pops, nx0, ny0)
        self.fine = OracleLevel(L, h, 1, L.orc_oct2_field, L.orc_oct2_pops, nx1.value, ny1.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_oct2_free(h)
            self._h = None

    def xi(self):
        out = np.zeros(3)
        self._L.orc_oct2_xi(self._h, _dp(out))
        return out

    def coarse_to_fine(self):
        self._L.orc_oct2_coarse_to_fine(self._h)

    def fine_to_coarse(self):
        self._L.orc_oct2_fine_to_coarse(self._h)

    def ghost_ks(self):
        return [self._L.orc_oct2_ghost_k(self._h, j)
                for j in range(self._L.orc_oct2_n_ghosts(self._h))]

    def coincident(self):
        out = []
        for j in range(self._L.orc_oct2_n_coincident(self._h)):
            k = C.c_size_t()
            kf = (C.c_size_t * 9)()
            w = np.zeros(9)
            self._L.orc_oct2_coincident(self._h, j, C.byref(k), kf, _dp(w))
            out.append((k.value, list(kf), w))
        return out


def oct_relax_times(nu, alpha):
    out = np.zeros(3)
    lib().orc_oct_relax_times(nu, alpha, _dp(out))
    return out


def oct_rescaling_factors(tau_c, tau_f, substeps=2, odd_exact=False):
    out = np.zeros(3)
    a = np.ascontiguousarray(tau_c, dtype=np.float64)
    b = np.ascontiguousarray(tau_f, dtype=np.float64)
    lib().orc_oct_rescaling_factors(_dp(a), _dp(b), substeps, int(odd_exact), _dp(out))
    return out


def oct_rescale_fneq(fneq, xi, inverse=False):
    a = np.ascontiguousarray(fneq, dtype=np.float64)
    x = np.ascontiguousarray(xi, dtype=np.float64)
    out = np.zeros(9)
    fn = lib().orc_oct_rescale_fneq_inverse if inverse else lib().orc_oct_rescale_fneq
    fn(_dp(a), _dp(x), _dp(out))
    return out


def oct_bilinear_stencil(x, y, nx0, ny0):
    k = (C.c_size_t * 4)()
    w = np.zeros(4)
    lib().orc_oct_bilinear_stencil(x, y, nx0, ny0, k, _dp(w))
    return list(k), w
