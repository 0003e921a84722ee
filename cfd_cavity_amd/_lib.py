"""ctypes loader for libcavlbm.so (the C-ABI declared in include/cavlbm.h).

The library is the only compute path: if it is missing or fails to load,
``lib()`` raises; nothing falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcavlbm.so")

E_PARAM, E_SIZE, E_HIP, E_UNSUPPORTED, E_NODEVICE, E_IO = -1, -2, -3, -4, -5, -6
STORAGE = {"fp64": 0, "fp32": 1, "fp16_offset": 2}
FIELD = {"rho": 0, "ux": 1, "uy": 2, "temperature": 3, "nu_t": 4, "fy_node": 5}
PHASE = {"collide": 0, "stream": 1, "thermal_bc": 2, "update_macroscopic": 3,
         "enforce_wall_velocity": 4, "bounce_back": 5}


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_uint64), ("ny", C.c_uint64),
        ("ra", C.c_double), ("pr", C.c_double), ("cs", C.c_double),
        ("pr_t", C.c_double), ("utarget", C.c_double),
        ("seed", C.c_uint64), ("jobs", C.c_uint64),
        ("alpha_mult", C.c_double), ("nu_start_mult", C.c_double),
        ("has_nu_override", C.c_int32), ("storage", C.c_int32),
        ("nu_override", C.c_double),
        ("geq_order", C.c_int32), ("device", C.c_int32),
    ]


class Diag(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "nusselt_hot", "nusselt_cold", "nusselt_hot_wall_link", "nusselt_volume",
        "total_mass", "total_f_mass", "max_uy", "max_u", "nu_t_max", "y_plus_hot")] + [
        ("all_finite", C.c_int32), ("reserved", C.c_int32), ("time_step", C.c_uint64)]


class TuneResult(C.Structure):
    _fields_ = [("tile_x", C.c_int32), ("tile_y", C.c_int32), ("xcd_remap", C.c_int32),
                ("candidates", C.c_int32), ("best_ms_per_step", C.c_double),
                ("first_ms_per_step", C.c_double), ("steps_per_launch", C.c_int32),
                ("reserved", C.c_int32)]


class Derived(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "lz", "lx", "nu", "alpha", "alpha_phys", "beta_eff", "tau", "tau_t", "cs",
        "pr_t", "nu_start_mult", "t_ref")] + [("spin_up_steps", C.c_uint64)]


class OctreeParams(C.Structure):
    _fields_ = [
        ("nx0", C.c_uint64), ("ny0", C.c_uint64), ("bl_width", C.c_uint64),
        ("ra", C.c_double), ("pr", C.c_double), ("utarget", C.c_double),
        ("seed", C.c_uint64), ("jobs", C.c_uint64),
        ("odd_exact", C.c_int32), ("levels", C.c_int32),
        ("nu_start_mult", C.c_double),
        ("has_nu_override", C.c_int32), ("geq_order", C.c_int32),
        ("nu_override", C.c_double),
        ("device", C.c_int32), ("reserved", C.c_int32),
    ]


class OctreeDerived(C.Structure):
    _fields_ = [("nu_lu", C.c_double), ("alpha_lu", C.c_double),
                ("beta_eff_coarse", C.c_double), ("beta_eff_fine", C.c_double),
                ("tau", (C.c_double * 3) * 3), ("xi_a", C.c_double * 3),
                ("xi_b", C.c_double * 3), ("spin_up_steps", C.c_uint64),
                ("n_nodes_total", C.c_uint64), ("levels", C.c_int32), ("n_grids", C.c_int32)]


class OctreeDiag(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "nusselt_hot", "nusselt_cold", "nusselt_hot_link0", "nusselt_hot_fine", "max_u_all",
        "max_rho_dev", "max_rho_dev_ifaces", "masse")] + [
        ("guard", C.c_int32), ("reserved", C.c_int32), ("time_step", C.c_uint64)]


# name -> (restype, argtypes); the exported set of include/cavlbm.h
_vp, _u64, _i32, _d = C.c_void_p, C.c_uint64, C.c_int32, C.c_double
_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "cavlbm_abi_version": (C.c_int, []),
    "cavlbm_params_default": (C.c_int, [C.POINTER(Params)]),
    "cavlbm_create": (C.c_int, [C.POINTER(Params), C.POINTER(_vp)]),
    "cavlbm_destroy": (C.c_int, [_vp]),
    "cavlbm_step": (C.c_int, [_vp]),
    "cavlbm_run": (C.c_int, [_vp, _u64]),
    "cavlbm_synchronize": (C.c_int, [_vp]),
    "cavlbm_init_from_fields": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _u64]),
    "cavlbm_get_field": (C.c_int, [_vp, _i32, _dp, _u64]),
    "cavlbm_set_field": (C.c_int, [_vp, _i32, _dp, _u64]),
    "cavlbm_get_pops": (C.c_int, [_vp, _i32, _dp, _u64]),
    "cavlbm_set_pops": (C.c_int, [_vp, _i32, _dp, _u64]),
    "cavlbm_diagnostics": (C.c_int, [_vp, C.POINTER(Diag)]),
    "cavlbm_max_uy_on_row": (C.c_int, [_vp, _u64, _dp, _dp]),
    "cavlbm_derived_params": (C.c_int, [_vp, C.POINTER(Derived)]),
    "cavlbm_time_step": (C.c_int, [_vp, C.POINTER(_u64)]),
    "cavlbm_set_time_step": (C.c_int, [_vp, _u64]),
    "cavlbm_set_hot_wall": (C.c_int, [_vp, _dp, _u64]),
    "cavlbm_set_dynamic_sgs": (C.c_int, [_vp, _i32, _d, _d, _d]),
    "cavlbm_dynamic_stats": (C.c_int, [_vp, _dp]),
    "cavlbm_get_sgs_pre": (C.c_int, [_vp, _i32, _dp, _u64]),
    "cavlbm_set_jobs": (C.c_int, [_u64]),
    "cavlbm_phase": (C.c_int, [_vp, _i32, _u64]),
    "cavlbm_time_run": (C.c_int, [_vp, _u64, _dp]),
    "cavlbm_time_step_kernel": (C.c_int, [_vp, _u64, _dp]),
    "cavlbm_set_time_blocking": (C.c_int, [_vp, _i32]),
    "cavlbm_steps_per_launch": (C.c_int, [_vp, C.POINTER(_i32)]),
    "cavlbm_set_dynamic_fusion": (C.c_int, [_vp, _i32]),
    "cavlbm_dynamic_launches": (C.c_int, [_vp, C.POINTER(_i32)]),
    "cavlbm_last_error": (C.c_char_p, [_vp]),
    "cavlbm_average_reset": (C.c_int, [_vp]),
    "cavlbm_average_accumulate": (C.c_int, [_vp]),
    "cavlbm_average_count": (C.c_int, [_vp, C.POINTER(_u64)]),
    "cavlbm_average_means": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, _u64]),
    "cavlbm_snapshot_write": (C.c_int, [C.c_char_p, _u64, _u64, _dp, _dp, _dp, _dp]),
    "cavlbm_snapshot_write_atomic": (C.c_int, [C.c_char_p, _u64, _u64, _dp, _dp, _dp, _dp]),
    "cavlbm_snapshot_read": (C.c_int, [C.c_char_p, C.POINTER(_u64), C.POINTER(_u64), _dp, _dp,
                                       _dp, _dp, _u64]),
    "cavlbm_interpolate_field": (C.c_int, [_dp, _u64, _u64, _dp, _u64, _u64]),
    "cavlbm_snapshot_last_error": (C.c_char_p, []),
    "cavlbm_nu_window": (C.c_int, [_vp, _dp, _u64, _d, _u64, _dp]),
    "cavlbm_nu_window_static": (C.c_int, [_vp, _dp, _u64, _d, _u64, _dp]),
    "cavlbm_adjoint_nu_window": (C.c_int, [_vp, _dp, _u64, _d, _u64, _u64, _dp, _dp, _dp]),
    "cavlbm_tangent_nu_window": (C.c_int, [_vp, _dp, _u64, _d, _dp, _d, _u64, _dp, _dp]),
    "cavlbm_batch_create": (C.c_int, [C.POINTER(Params), _u64, C.POINTER(_vp)]),
    "cavlbm_batch_destroy": (C.c_int, [_vp]),
    "cavlbm_batch_run": (C.c_int, [_vp, _u64]),
    "cavlbm_batch_synchronize": (C.c_int, [_vp]),
    "cavlbm_batch_solver": (C.c_int, [_vp, _u64, C.POINTER(_vp)]),
    "cavlbm_batch_autotune": (C.c_int, [_vp, _u64, C.POINTER(TuneResult)]),
    "cavlbm_batch_nodes": (C.c_int, [_vp, C.POINTER(_u64)]),
    "cavlbm_batch_time_run": (C.c_int, [_vp, _u64, _dp]),
    "cavlbm_batch_last_error": (C.c_char_p, [_vp]),
    "cavlbm_octree_params_default": (C.c_int, [C.POINTER(OctreeParams)]),
    "cavlbm_octree_create": (C.c_int, [C.POINTER(OctreeParams), C.POINTER(_vp)]),
    "cavlbm_octree_destroy": (C.c_int, [_vp]),
    "cavlbm_octree_step": (C.c_int, [_vp]),
    "cavlbm_octree_run": (C.c_int, [_vp, _u64]),
    "cavlbm_octree_synchronize": (C.c_int, [_vp]),
    "cavlbm_octree_time_run": (C.c_int, [_vp, _u64, _dp]),
    "cavlbm_octree_time_step": (C.c_int, [_vp, C.POINTER(_u64)]),
    "cavlbm_octree_derived_params": (C.c_int, [_vp, C.POINTER(OctreeDerived)]),
    "cavlbm_octree_grid_dims": (C.c_int, [_vp, _i32, C.POINTER(_u64), C.POINTER(_u64)]),
    "cavlbm_octree_get_field": (C.c_int, [_vp, _i32, _i32, _dp, _u64]),
    "cavlbm_octree_set_field": (C.c_int, [_vp, _i32, _i32, _dp, _u64]),
    "cavlbm_octree_get_pops": (C.c_int, [_vp, _i32, _i32, _dp, _u64]),
    "cavlbm_octree_set_pops": (C.c_int, [_vp, _i32, _i32, _dp, _u64]),
    "cavlbm_octree_set_from_snapshot": (C.c_int, [_vp, _dp, _dp, _dp, _dp, _u64]),
    "cavlbm_octree_diagnostics": (C.c_int, [_vp, C.POINTER(OctreeDiag)]),
    "cavlbm_octree_set_dynamic_sgs": (C.c_int, [_vp, _i32, _d, _d, _d]),
    "cavlbm_octree_dynamic_stats": (C.c_int, [_vp, _dp]),
    "cavlbm_octree_get_sgs_pre": (C.c_int, [_vp, _i32, _i32, _dp, _u64]),
    "cavlbm_octree_interface_factors": (C.c_int, [_vp, _dp]),
    "cavlbm_octree_last_error": (C.c_char_p, [_vp]),
    "cavlbm_octsnap_write": (C.c_int, [C.c_char_p, C.c_uint32, _u64, C.c_uint32,
                                       C.POINTER(_u64), _dp, _dp, _dp, _dp, _i32]),
    "cavlbm_octsnap_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(_u64),
                                      C.POINTER(C.c_uint32), C.POINTER(_u64), _dp, _dp, _dp,
                                      _dp, _u64]),
}

_lib = None


class CavlbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cavlbm error {code}: {msg}")
        self.code = code


def lib():
    """Load libcavlbm.so once; raise if it is missing (no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int, handle=None) -> None:
    if rc != 0:
        msg = lib().cavlbm_last_error(handle)
        raise CavlbmError(rc, msg.decode() if msg else "")
