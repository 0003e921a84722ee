"""Host-side mirror of the reference's ``CavitySolver`` over the C-ABI.

Same constructor arguments, method names, argument meaning and error
behaviour as ``crate::uniform::solver::CavitySolver``
(bbrangeo/cfd_cavity src/uniform/solver.rs:35-104, :108-293, :1012-1020),
with the arithmetic running in libcavlbm's HIP kernels on a gfx950 GPU.
Where the reference panics on an ``assert!``, this raises ``ValueError``.

Field reads return numpy copies; the reference's direct field writes
(``s.lbm.temperature[k] = ...``) become ``set_field``/``set_pops``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

T_HOT = 0.5          # src/uniform/solver.rs:17
T_COLD = -0.5        # :19
DELTA_T = T_HOT - T_COLD  # :21


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(nx, ny, ra, pr, cs, pr_t, utarget, seed, jobs=1, alpha_mult=1.0,
                nu_start_mult=1.0, nu_override=None, *, storage="fp64", geq_order=2, device=0):
    """The C-ABI parameter struct for CavitySolver::new's twelve arguments."""
    p = _lib.Params()
    check(lib().cavlbm_params_default(C.byref(p)))
    p.nx, p.ny = int(nx), int(ny)
    p.ra, p.pr, p.cs, p.pr_t, p.utarget = ra, pr, cs, pr_t, utarget
    p.seed, p.jobs = int(seed), int(jobs)
    p.alpha_mult, p.nu_start_mult = alpha_mult, nu_start_mult
    p.has_nu_override = 0 if nu_override is None else 1
    p.nu_override = 0.0 if nu_override is None else float(nu_override)
    p.storage = _lib.STORAGE[storage]
    p.geq_order = int(geq_order)
    p.device = int(device)
    return p


class _LbmView:
    """Read-only view mirroring the reference's ``s.lbm`` (ThermalD2Q9) fields."""

    def __init__(self, s: "CavitySolver"):
        self._s = s

    @property
    def nx(self):
        return self._s.nx()

    @property
    def ny(self):
        return self._s.ny()

    @property
    def rho(self):
        return self._s.get_field("rho")

    @property
    def ux(self):
        return self._s.get_field("ux")

    @property
    def uy(self):
        return self._s.get_field("uy")

    @property
    def temperature(self):
        return self._s.get_field("temperature")

    @property
    def f(self):
        return self._s.get_pops("f")

    @property
    def g(self):
        return self._s.get_pops("g")

    @property
    def tau(self):
        return self._s.derived()["tau"]

    @property
    def tau_t(self):
        return self._s.derived()["tau_t"]

    @property
    def t_ref(self):
        return self._s.derived()["t_ref"]


class CavitySolver:
    """``CavitySolver::new(nx, ny, ra, pr, cs, pr_t, utarget, seed, jobs,
    alpha_mult, nu_start_mult, nu_override)`` (src/uniform/solver.rs:108).

    Extra keyword options: ``storage`` ("fp64", "fp32", "fp16_offset"),
    ``geq_order`` (2 quadratic, 1 linear g equilibrium) and ``device``.
    """

    def __init__(self, nx, ny, ra, pr, cs, pr_t, utarget, seed, jobs=1,
                 alpha_mult=1.0, nu_start_mult=1.0, nu_override=None, *,
                 storage="fp64", geq_order=2, device=0):
        L = lib()
        p = make_params(nx, ny, ra, pr, cs, pr_t, utarget, seed, jobs, alpha_mult,
                        nu_start_mult, nu_override, storage=storage, geq_order=geq_order,
                        device=device)
        h = C.c_void_p()
        rc = L.cavlbm_create(C.byref(p), C.byref(h))
        if rc == _lib.E_PARAM:
            raise ValueError(L.cavlbm_last_error(None).decode())
        check(rc, None)
        self._adopt(h, int(nx), int(ny), storage, owned=True)
        self.jobs = max(int(jobs), 1)

    def _adopt(self, h, nx, ny, storage, owned):
        self._h = h
        self._L = lib()
        self._nx, self._ny = nx, ny
        self.storage = storage
        self._owned = owned
        self._tblock = None
        self.lbm = _LbmView(self)

    @classmethod
    def _borrow(cls, handle, nx, ny, storage, keepalive):
        """Wrap a handle owned elsewhere (a batch cavity); never destroyed here."""
        obj = cls.__new__(cls)
        obj._adopt(handle, nx, ny, storage, owned=False)
        obj.jobs = 1
        obj._keepalive = keepalive
        return obj

    # -- lifetime ---------------------------------------------------------
    def close(self):
        h = getattr(self, "_h", None)
        if h and getattr(self, "_owned", False):
            self._L.cavlbm_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        check(rc, self._h)

    # -- reference API ----------------------------------------------------
    def nx(self):
        return self._nx

    def ny(self):
        return self._ny

    def idx(self, ix, iy):
        return iy * self._nx + ix

    def step(self):
        self._check(self._L.cavlbm_step(self._h))

    def run(self, nsteps):
        self._check(self._L.cavlbm_run(self._h, int(nsteps)))

    def synchronize(self):
        self._check(self._L.cavlbm_synchronize(self._h))

    @staticmethod
    def set_jobs(jobs):
        lib().cavlbm_set_jobs(int(jobs))

    def set_dynamic_sgs(self, dynamic_smag, smag_alpha=0.5, nu_sm_min=-0.1, eta_test=2.0):
        """CavitySolver::set_dynamic_sgs (solver.rs:215-220)."""
        self._check(self._L.cavlbm_set_dynamic_sgs(self._h, 1 if dynamic_smag else 0,
                                                    smag_alpha, nu_sm_min, eta_test))
        self._dyn = (bool(dynamic_smag), smag_alpha, nu_sm_min, eta_test)

    @property
    def dynamic_smag(self):
        return getattr(self, "_dyn", (False,))[0]

    def dynamic_stats(self):
        """(cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max) of the last
        dynamic step (solver.rs:821-824); zeros before the first."""
        out = np.zeros(4)
        self._check(self._L.cavlbm_dynamic_stats(self._h, _dp(out)))
        return tuple(float(v) for v in out)

    @property
    def cs_dyn_mean(self):
        return self.dynamic_stats()[0]

    @property
    def cs_dyn_max(self):
        return self.dynamic_stats()[1]

    @property
    def kappa_dyn_mean(self):
        return self.dynamic_stats()[2]

    @property
    def kappa_dyn_max(self):
        return self.dynamic_stats()[3]

    def _sgs_pre(self, which):
        n = self._nx * self._ny
        out = np.empty(n)
        rc = self._L.cavlbm_get_sgs_pre(self._h, which, _dp(out), n)
        if rc == _lib.E_PARAM:
            return None  # the reference's None before the first dynamic step
        self._check(rc)
        return out

    @property
    def nu_t_pre(self):
        """Per-node nu_t of the last dynamic step, or None (solver.rs:89)."""
        return self._sgs_pre(0)

    @property
    def alpha_t_pre(self):
        """Per-node alpha_t of the last dynamic step, or None (solver.rs:93)."""
        return self._sgs_pre(1)

    def init_from_fields(self, rho, ux, uy, t):
        n = self._nx * self._ny
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (rho, ux, uy, t)]
        if any(a.size != n for a in arrs):
            raise ValueError(
                f"init_from_fields: fields of size {n} expected "
                f"(rho={arrs[0].size}, ux={arrs[1].size}, uy={arrs[2].size}, t={arrs[3].size})")
        self._check(self._L.cavlbm_init_from_fields(self._h, *[_dp(a) for a in arrs], n))

    def total_mass(self):
        return self.diagnostics()["total_mass"]

    def total_f_mass(self):
        return self.diagnostics()["total_f_mass"]

    @property
    def time_step(self):
        v = C.c_uint64()
        self._check(self._L.cavlbm_time_step(self._h, C.byref(v)))
        return v.value

    @time_step.setter
    def time_step(self, t):
        self._check(self._L.cavlbm_set_time_step(self._h, int(t)))

    def derived(self):
        d = _lib.Derived()
        self._check(self._L.cavlbm_derived_params(self._h, C.byref(d)))
        return {name: getattr(d, name) for name, _ in d._fields_}

    # public scalars of the Rust struct
    @property
    def lz(self):
        return self.derived()["lz"]

    @property
    def lx(self):
        return self.derived()["lx"]

    @property
    def beta_eff(self):
        return self.derived()["beta_eff"]

    @property
    def nu(self):
        return self.derived()["nu"]

    @property
    def alpha(self):
        return self.derived()["alpha"]

    @property
    def alpha_phys(self):
        return self.derived()["alpha_phys"]

    @property
    def cs(self):
        return self.derived()["cs"]

    @property
    def pr_t(self):
        return self.derived()["pr_t"]

    @property
    def spin_up_steps(self):
        return self.derived()["spin_up_steps"]

    @property
    def nu_t(self):
        return self.get_field("nu_t")

    @property
    def fy_node(self):
        return self.get_field("fy_node")

    # -- state exchange ---------------------------------------------------
    def get_field(self, name):
        n = self._nx * self._ny
        out = np.empty(n, dtype=np.float64)
        self._check(self._L.cavlbm_get_field(self._h, _lib.FIELD[name], _dp(out), n))
        return out

    def set_field(self, name, values):
        n = self._nx * self._ny
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if a.size != n:
            raise ValueError(f"set_field: {n} values expected, got {a.size}")
        self._check(self._L.cavlbm_set_field(self._h, _lib.FIELD[name], _dp(a), n))

    def get_pops(self, which):
        n = self._nx * self._ny
        out = np.empty((n, 9), dtype=np.float64)
        self._check(self._L.cavlbm_get_pops(self._h, 0 if which == "f" else 1, _dp(out), 9 * n))
        return out

    def set_pops(self, which, values):
        n = self._nx * self._ny
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(n * 9)
        self._check(self._L.cavlbm_set_pops(self._h, 0 if which == "f" else 1, _dp(a), 9 * n))

    def set_hot_wall(self, theta_hot):
        if theta_hot is None:
            self._check(self._L.cavlbm_set_hot_wall(self._h, None, 0))
            return
        a = np.ascontiguousarray(theta_hot, dtype=np.float64).ravel()
        self._check(self._L.cavlbm_set_hot_wall(self._h, _dp(a), a.size))

    # -- crate-private phases (tests call them directly in the reference) --
    def phase(self, name, step=0):
        self._check(self._L.cavlbm_phase(self._h, _lib.PHASE[name], int(step)))

    def collide(self, step):
        self.phase("collide", step)

    def stream(self):
        self.phase("stream")

    def apply_thermal_boundary_conditions(self):
        self.phase("thermal_bc")

    def update_macroscopic(self):
        self.phase("update_macroscopic")

    def enforce_wall_velocity(self):
        self.phase("enforce_wall_velocity")

    def apply_bounce_back(self):
        self.phase("bounce_back")

    # -- diagnostics --------------------------------------------------------
    def diagnostics(self):
        d = _lib.Diag()
        self._check(self._L.cavlbm_diagnostics(self._h, C.byref(d)))
        out = {name: getattr(d, name) for name, _ in d._fields_ if name != "reserved"}
        out["all_finite"] = bool(out["all_finite"])
        return out

    def max_uy_on_row(self, iy):
        w = C.c_double()
        x = C.c_double()
        self._check(self._L.cavlbm_max_uy_on_row(self._h, int(iy), C.byref(w), C.byref(x)))
        return w.value, x.value

    # -- time averaging (TimeAverager, src/uniform/post.rs:11-57) ------------
    def average_reset(self):
        self._check(self._L.cavlbm_average_reset(self._h))

    def average_accumulate(self):
        self._check(self._L.cavlbm_average_accumulate(self._h))

    def average_count(self):
        v = C.c_uint64()
        self._check(self._L.cavlbm_average_count(self._h, C.byref(v)))
        return v.value

    def average_means(self):
        """(ux, uy, T, nu_t, u^2) means, like TimeAverager::means."""
        n = self._nx * self._ny
        out = [np.empty(n) for _ in range(5)]
        self._check(self._L.cavlbm_average_means(self._h, *[_dp(a) for a in out], n))
        return tuple(out)

    # -- timing helpers -----------------------------------------------------
    def time_run(self, nsteps):
        ms = C.c_double()
        self._check(self._L.cavlbm_time_run(self._h, int(nsteps), C.byref(ms)))
        return ms.value

    def time_step_kernel(self, launches):
        """Mean duration (ms) of one launch of the run loop's step kernel
        (two steps per launch when ``steps_per_launch`` is 2)."""
        ms = C.c_double()
        self._check(self._L.cavlbm_time_step_kernel(self._h, int(launches), C.byref(ms)))
        return ms.value

    @property
    def time_blocking(self):
        """Most steps per launch requested for the fused run loop (1..4), or
        None for the library's default (FP64 4, FP32 and FP16-offset 3)."""
        return self._tblock

    @time_blocking.setter
    def time_blocking(self, steps):
        self._check(self._L.cavlbm_set_time_blocking(self._h, int(steps)))
        self._tblock = int(steps)

    @property
    def steps_per_launch(self):
        """Steps per launch the run loop uses for this lattice and storage."""
        out = C.c_int32()
        self._check(self._L.cavlbm_steps_per_launch(self._h, C.byref(out)))
        return out.value

    def set_dynamic_fusion(self, one_launch):
        """One launch per dynamic-SGS step (True, the default, where it applies:
        a test filter of 1..3 nodes, any storage) or four (False); both give
        the same state bit for bit."""
        self._check(self._L.cavlbm_set_dynamic_fusion(self._h, 1 if one_launch else 0))

    @property
    def dynamic_launches(self):
        """Kernel launches per dynamic-SGS step for the current settings."""
        out = C.c_int32()
        self._check(self._L.cavlbm_dynamic_launches(self._h, C.byref(out)))
        return out.value
