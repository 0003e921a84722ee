"""Host-side mirror of the reference's octree ``CavitySolver``
(bbrangeo/cfd_cavity src/octree/solver.rs:118-1206) over the C-ABI of
libcavlbm (``cavlbm_octree_*`` in include/cavlbm.h).

Levels: grid 0 is the coarse base grid L0 over the whole cavity, grids 1 / 2
the hot / cold wall blocks L1 (dx/2), and with ``levels=3`` grids 3 / 4 the L2
strips (dx/4) inside them -- the order of the reference's snapshots
(solver.rs:1048-1066).  Every step runs on the GPU; the constructor's asserts
raise ValueError like the reference's panics.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import CavlbmError, lib

FIELDS = ("rho", "ux", "uy", "temperature", "fy_node")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OctreeCavitySolver:
    """``CavitySolver::new(nx0, ny0, bl_width, ra, pr, utarget, seed, jobs,
    odd_exact, levels, nu_start_mult, nu_override)`` (solver.rs:340-353)."""

    def __init__(self, nx0, ny0, bl_width, ra, pr=0.71, utarget=0.02, seed=42, jobs=1,
                 odd_exact=False, levels=2, nu_start_mult=1.0, nu_override=None, *,
                 geq_order=2, device=0):
        L = lib()
        p = _lib.OctreeParams()
        L.cavlbm_octree_params_default(C.byref(p))
        p.nx0, p.ny0, p.bl_width = nx0, ny0, bl_width
        p.ra, p.pr, p.utarget = ra, pr, utarget
        p.seed, p.jobs = seed, jobs
        p.odd_exact, p.levels = int(bool(odd_exact)), levels
        p.nu_start_mult = nu_start_mult
        p.has_nu_override = 0 if nu_override is None else 1
        p.nu_override = 0.0 if nu_override is None else nu_override
        p.geq_order, p.device = geq_order, device
        h = C.c_void_p()
        rc = L.cavlbm_octree_create(C.byref(p), C.byref(h))
        if rc != 0:
            msg = (L.cavlbm_octree_last_error(None) or b"").decode()
            if rc == _lib.E_PARAM:
                raise ValueError(msg)
            raise CavlbmError(rc, msg)
        self._h, self._L = h, L
        self.nx0, self.ny0, self.levels = nx0, ny0, levels
        d = self.derived()
        self.n_grids = d["n_grids"]
        self.dims = []
        for gi in range(self.n_grids):
            nx, ny = C.c_uint64(), C.c_uint64()
            self._check(L.cavlbm_octree_grid_dims(h, gi, C.byref(nx), C.byref(ny)))
            self.dims.append((nx.value, ny.value))

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.cavlbm_octree_destroy(h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            msg = self._L.cavlbm_octree_last_error(self._h)
            raise CavlbmError(rc, msg.decode() if msg else "")

    # --- stepping (solver.rs:647-947)
    def step_cavity(self):
        self._check(self._L.cavlbm_octree_step(self._h))

    step = step_cavity

    def run(self, nsteps):
        self._check(self._L.cavlbm_octree_run(self._h, int(nsteps)))

    def synchronize(self):
        self._check(self._L.cavlbm_octree_synchronize(self._h))

    def time_run(self, nsteps):
        ms = C.c_double()
        self._check(self._L.cavlbm_octree_time_run(self._h, int(nsteps), C.byref(ms)))
        return ms.value

    @property
    def time_step(self):
        t = C.c_uint64()
        self._check(self._L.cavlbm_octree_time_step(self._h, C.byref(t)))
        return t.value

    def set_dynamic_sgs(self, on, smag_alpha=0.5, nu_sm_min=-0.1, eta_test=2.0):
        """solver.rs:475-480: the dynamic mixed-scale model on every level
        (src/octree/smag.rs) before each of its collisions."""
        self._check(self._L.cavlbm_octree_set_dynamic_sgs(self._h, int(bool(on)), smag_alpha,
                                                          nu_sm_min, eta_test))

    def dynamic_stats(self):
        """(cs_dyn_mean, cs_dyn_max, kappa_dyn_mean, kappa_dyn_max, nu_t_max)."""
        out = np.zeros(5)
        self._check(self._L.cavlbm_octree_dynamic_stats(self._h, _dp(out)))
        return tuple(float(v) for v in out)

    def sgs_pre(self, grid, which=0):
        """nu_t_pre (0) or alpha_t_pre (1) of a grid."""
        nx, ny = self.dims[grid]
        out = np.empty(nx * ny)
        self._check(self._L.cavlbm_octree_get_sgs_pre(self._h, grid, which, _dp(out), out.size))
        return out

    def interface_factors(self):
        """xi (even, odd, thermal) of interface A, then B of the hot and cold block."""
        out = np.zeros(9)
        self._check(self._L.cavlbm_octree_interface_factors(self._h, _dp(out)))
        return out

    # --- state
    def derived(self):
        d = _lib.OctreeDerived()
        self._check(self._L.cavlbm_octree_derived_params(self._h, C.byref(d)))
        out = {k: getattr(d, k) for k in ("nu_lu", "alpha_lu", "beta_eff_coarse",
                                           "beta_eff_fine", "spin_up_steps", "n_nodes_total",
                                           "levels", "n_grids")}
        out["tau"] = [list(row) for row in d.tau]
        out["xi_a"], out["xi_b"] = list(d.xi_a), list(d.xi_b)
        return out

    def get_field(self, grid, name):
        nx, ny = self.dims[grid]
        out = np.empty(nx * ny)
        self._check(self._L.cavlbm_octree_get_field(self._h, grid, _lib.FIELD[name], _dp(out),
                                                    out.size))
        return out

    def set_field(self, grid, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        self._check(self._L.cavlbm_octree_set_field(self._h, grid, _lib.FIELD[name], _dp(a),
                                                    a.size))

    def get_pops(self, grid, which):
        nx, ny = self.dims[grid]
        out = np.empty((nx * ny, 9))
        self._check(self._L.cavlbm_octree_get_pops(self._h, grid, 0 if which == "f" else 1,
                                                   _dp(out), out.size))
        return out

    def set_pops(self, grid, which, values):
        a = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self._L.cavlbm_octree_set_pops(self._h, grid, 0 if which == "f" else 1,
                                                   _dp(a), a.size))

    def level_fields(self):
        """(nx, ny, rho, ux, uy, T) of every grid in snapshot order
        (solver.rs:1048-1066)."""
        return [(nx, ny) + tuple(self.get_field(gi, n) for n in FIELDS[:4])
                for gi, (nx, ny) in enumerate(self.dims)]

    def set_from_snapshot(self, rho, ux, uy, t):
        """solver.rs:1095-1175 with every grid's fields concatenated."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (rho, ux, uy, t)]
        self._check(self._L.cavlbm_octree_set_from_snapshot(self._h, *[_dp(a) for a in arrs],
                                                            arrs[0].size))

    def write_snapshot(self, path, step, atomic=True):
        """OCTSNAP1 export of the level fields (solver.rs:1068-1075,
        snapshot.rs:140-161)."""
        from .snapshot import write_octree_snapshot

        write_octree_snapshot(path, self.levels, step, self.level_fields(), atomic)

    def restart(self, path):
        """--restart of run_octree (main.rs:526-546): read an OCTSNAP1 file and
        inject its fields (set_from_snapshot, which requires the same layout,
        solver.rs:1095-1133).  Returns the snapshot's step."""
        from .snapshot import read_octree_snapshot

        levels, step, grids = read_octree_snapshot(path)
        n_expected = 3 if self.levels == 2 else 5
        if levels != self.levels or len(grids) != n_expected:
            raise ValueError(
                f"restart : snapshot levels={levels} ({len(grids)} grilles) incompatible avec "
                f"le solveur levels={self.levels} ({n_expected} grilles attendues)")
        for gi, (g, d) in enumerate(zip(grids, self.dims)):
            if tuple(g[:2]) != tuple(d):
                raise ValueError(
                    f"restart : grille {gi} du snapshot {g[0]}x{g[1]} ≠ solveur {d[0]}x{d[1]} "
                    "(même layout requis — pas d'interpolation inter-grille en Étape G)")
        self.set_from_snapshot(*(np.concatenate([g[2 + q] for g in grids]) for q in range(4)))
        return step

    # --- diagnostics (cavity.rs:672-790, solver.rs:949-1041)
    def diagnostics(self):
        d = _lib.OctreeDiag()
        self._check(self._L.cavlbm_octree_diagnostics(self._h, C.byref(d)))
        return {k: getattr(d, k) for k, _ in _lib.OctreeDiag._fields_ if k != "reserved"}

    def noise_stats(self):
        d = self.diagnostics()
        return dict(max_rho_dev=d["max_rho_dev"], max_rho_dev_ifaces=d["max_rho_dev_ifaces"],
                    masse=d["masse"])

    def max_u_all(self):
        return self.diagnostics()["max_u_all"]

    def check_guards(self):
        """None, or the reference's abort message (solver.rs:1018-1041)."""
        d = self.diagnostics()
        if d["guard"] == 1:
            u = d["max_u_all"]
            return (f"AVORT : max_u = {u:.5f} dépasse le garde-fou 0.10 "
                    f"(Ma_local = max_u·√3 ≈ {u * 3 ** 0.5:.3f})")
        if d["guard"] == 2:
            return "DIVERGENCE détectée (valeurs non finies)."
        return None
