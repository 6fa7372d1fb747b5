"""ctypes binding of libclarabel_hip.so (include/clarabel_hip.h).

The package only ever binds the in-tree HIP build; when it is missing the
import of the solver fails with an error that says how to build it.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(HERE, "libclarabel_hip.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)


class ClhSchedule(C.Structure):
    _fields_ = [
        ("relax_max_cols", C.c_int32), ("relax_zero_pct", C.c_int32), ("group_lanes", C.c_int32),
        ("blocks_per_cu", C.c_int32), ("grid_blocks", C.c_int32), ("tile_min_cols", C.c_int32),
        ("dense_min_cols", C.c_int32), ("use_tuned", C.c_int32),
    ]


class ClhTuneReport(C.Structure):
    _fields_ = [
        ("default_rate", C.c_double), ("tuned_rate", C.c_double), ("trials", C.c_int32), ("classes", C.c_int32),
        ("seconds", C.c_double),
    ]


class ClhSettings(C.Structure):
    _fields_ = [
        ("max_iter", C.c_int32), ("time_limit", C.c_double), ("verbose", C.c_int32),
        ("max_step_fraction", C.c_double),
        ("tol_gap_abs", C.c_double), ("tol_gap_rel", C.c_double), ("tol_feas", C.c_double),
        ("tol_infeas_abs", C.c_double), ("tol_infeas_rel", C.c_double), ("tol_ktratio", C.c_double),
        ("reduced_tol_gap_abs", C.c_double), ("reduced_tol_gap_rel", C.c_double),
        ("reduced_tol_feas", C.c_double), ("reduced_tol_infeas_abs", C.c_double),
        ("reduced_tol_infeas_rel", C.c_double), ("reduced_tol_ktratio", C.c_double),
        ("equilibrate_enable", C.c_int32), ("equilibrate_max_iter", C.c_int32),
        ("equilibrate_min_scaling", C.c_double), ("equilibrate_max_scaling", C.c_double),
        ("linesearch_backtrack_step", C.c_double), ("min_switch_step_length", C.c_double),
        ("min_terminate_step_length", C.c_double),
        ("static_regularization_enable", C.c_int32), ("static_regularization_constant", C.c_double),
        ("static_regularization_proportional", C.c_double),
        ("dynamic_regularization_enable", C.c_int32), ("dynamic_regularization_eps", C.c_double),
        ("dynamic_regularization_delta", C.c_double),
        ("iterative_refinement_enable", C.c_int32), ("iterative_refinement_reltol", C.c_double),
        ("iterative_refinement_abstol", C.c_double), ("iterative_refinement_max_iter", C.c_int32),
        ("iterative_refinement_stop_ratio", C.c_double),
        ("direct_solve_method", C.c_int32), ("soc_threshold", C.c_int32),
        ("mixed_switch_reltol", C.c_double), ("mixed_fp64_static_reg", C.c_double),
        ("serve_engine", C.c_int32), ("dynamic_regularization_keep_magnitude", C.c_int32),
        ("schedule", ClhSchedule),
    ]


class ClhInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("iterations", C.c_int32),
        ("obj_val", C.c_double), ("obj_val_dual", C.c_double),
        ("r_prim", C.c_double), ("r_dual", C.c_double),
        ("solve_time", C.c_double), ("setup_time", C.c_double),
        ("gap_abs", C.c_double), ("gap_rel", C.c_double), ("ktratio", C.c_double), ("mu", C.c_double),
        ("ir_steps", C.c_int32), ("kkt_solves", C.c_int32),
        ("nnz_L", C.c_int64), ("factor_flops", C.c_double),
        ("fp32_factors", C.c_int32), ("switched_fp64", C.c_int32),
        ("nnz_K", C.c_int64), ("kkt_dim", C.c_int32), ("n_supernodes", C.c_int32), ("n_levels", C.c_int32),
        ("pad", C.c_int32), ("phase_seconds", C.c_double * 8),
    ]


class ClhServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
        ("active_slots", C.c_int32), ("queued", C.c_int32),
        ("hbm_in_use", C.c_int64),
        ("p50_latency_s", C.c_double), ("p99_latency_s", C.c_double), ("mean_latency_s", C.c_double),
        ("busy_time_s", C.c_double),
        ("launches", C.c_int64),
        ("kernel_time_s", C.c_double), ("cancelled", C.c_int64),
    ]


class ClhLdlInfo(C.Structure):
    _fields_ = [
        ("nnz_L", C.c_int64), ("factor_flops", C.c_double),
        ("n_supernodes", C.c_int32), ("n_levels", C.c_int32),
        ("n_positive", C.c_int32), ("n_negative", C.c_int32),
        ("last_dynreg", C.c_int32), ("last_ir_steps", C.c_int32),
        ("last_refine_error", C.c_double), ("grid_blocks", C.c_int32), ("pad_ldl", C.c_int32),
    ]


class ClhConesInfo(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("degree", C.c_int32),
        ("n_zero", C.c_int32), ("n_nonneg", C.c_int32), ("n_soc", C.c_int32), ("n_exp", C.c_int32),
        ("n_pow", C.c_int32), ("n_psd", C.c_int32),
        ("symmetric", C.c_int32), ("pad", C.c_int32),
        ("hs_len", C.c_int64),
    ]


# every symbol include/clarabel_hip.h declares
EXPORTS = [
    "clh_last_error", "clh_version", "clh_device_info", "clh_default_settings",
    "clh_solver_create", "clh_solve", "clh_get_solution", "clh_get_trace",
    "clh_update_P", "clh_update_q", "clh_update_A", "clh_update_b", "clh_solver_destroy",
    "clh_solver_create_device", "clh_get_solution_device", "clh_update_P_device", "clh_update_q_device",
    "clh_update_A_device", "clh_update_b_device",
    "clh_server_create", "clh_server_submit", "clh_server_progress", "clh_server_progress_async",
    "clh_server_poll", "clh_server_wait",
    "clh_server_cancel",
    "clh_server_result", "clh_server_stats_get", "clh_server_destroy",
    "clh_server_autotune", "clh_solver_autotune", "clh_server_get_schedule", "clh_solver_get_schedule",
    "clh_ldl_create", "clh_ldl_refactor", "clh_ldl_solve", "clh_ldl_solve_refine", "clh_ldl_get_info",
    "clh_ldl_destroy", "clh_device_alloc", "clh_device_free", "clh_memcpy_to_device", "clh_memcpy_to_host",
    "clh_cones_create", "clh_cones_get_info", "clh_cones_update_scaling", "clh_cones_set_identity_scaling",
    "clh_cones_get_Hs", "clh_cones_mul_Hs", "clh_cones_affine_ds", "clh_cones_combined_ds_shift",
    "clh_cones_ds_from_dz_offset", "clh_cones_step_length", "clh_cones_step_length_noalloc",
    "clh_cones_compute_barrier", "clh_cones_margins",
    "clh_cones_scaled_unit_shift", "clh_cones_unit_initialization", "clh_cones_rectify_equilibration",
    "clh_cones_destroy",
]


def bind(path: str):
    """Load a library implementing the C ABI and declare its signatures."""
    L = C.CDLL(path)
    vp = C.c_void_p
    L.clh_last_error.restype = C.c_char_p
    L.clh_version.restype = C.c_char_p
    L.clh_device_info.argtypes = [C.c_char_p, C.c_size_t, ip]
    L.clh_default_settings.argtypes = [C.POINTER(ClhSettings)]
    L.clh_solver_create.argtypes = [C.POINTER(vp), C.c_int32, C.c_int32, ip, ip, dp, dp, ip, ip, dp, dp, C.c_int32,
                                    ip, ip, dp, C.POINTER(ClhSettings)]
    L.clh_solve.argtypes = [vp, vp]
    L.clh_get_solution.argtypes = [vp, dp, dp, dp, C.POINTER(ClhInfo)]
    L.clh_get_trace.argtypes = [vp, dp, C.c_int32, ip]
    for f in ("clh_update_P", "clh_update_q", "clh_update_A", "clh_update_b"):
        getattr(L, f).argtypes = [vp, dp]
    L.clh_solver_destroy.argtypes = [vp]
    L.clh_solver_destroy.restype = None
    L.clh_solver_create_device.argtypes = [C.POINTER(vp), C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp,
                                           C.c_int32, ip, ip, dp, C.POINTER(ClhSettings)]
    L.clh_get_solution_device.argtypes = [vp, vp, vp, vp, vp]
    for f in ("clh_update_P_device", "clh_update_q_device", "clh_update_A_device", "clh_update_b_device"):
        getattr(L, f).argtypes = [vp, vp]
    L.clh_server_create.argtypes = [C.POINTER(vp), C.c_int64, C.c_int32, C.POINTER(ClhSettings), vp]
    L.clh_server_submit.argtypes = [vp, C.POINTER(C.c_int64), C.c_int32, C.c_int32, ip, ip, dp, dp, ip, ip, dp, dp,
                                    C.c_int32, ip, ip, dp]
    L.clh_server_progress.argtypes = [vp, C.c_int32]
    L.clh_server_progress_async.argtypes = [vp, C.c_int32]
    L.clh_server_poll.argtypes = [vp, C.c_int64]
    L.clh_server_wait.argtypes = [vp, C.c_int64, C.c_int32, C.c_double]
    L.clh_server_cancel.argtypes = [vp, C.c_int64]
    L.clh_server_result.argtypes = [vp, C.c_int64, dp, dp, dp, C.POINTER(ClhInfo), dp]
    L.clh_server_stats_get.argtypes = [vp, C.POINTER(ClhServerStats)]
    L.clh_server_destroy.argtypes = [vp]
    L.clh_server_destroy.restype = None
    L.clh_server_autotune.argtypes = [vp, C.c_double, C.POINTER(ClhTuneReport)]
    L.clh_solver_autotune.argtypes = [vp, C.c_double, C.POINTER(ClhTuneReport)]
    L.clh_server_get_schedule.argtypes = [vp, C.c_int32, C.POINTER(ClhSchedule)]
    L.clh_solver_get_schedule.argtypes = [vp, C.POINTER(ClhSchedule)]
    L.clh_ldl_create.argtypes = [C.POINTER(vp), C.c_int32, ip, ip, C.c_int32, ip, C.c_int32]
    L.clh_ldl_refactor.argtypes = [vp, vp, C.c_double, C.c_int32, C.c_double, C.c_double, vp]
    L.clh_ldl_solve.argtypes = [vp, vp, vp, vp]
    L.clh_ldl_solve_refine.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double, C.c_int32, C.c_double, vp]
    L.clh_ldl_get_info.argtypes = [vp, C.POINTER(ClhLdlInfo)]
    L.clh_ldl_destroy.argtypes = [vp]
    L.clh_ldl_destroy.restype = None
    L.clh_device_alloc.argtypes = [C.POINTER(vp), C.c_size_t]
    L.clh_device_free.argtypes = [vp]
    L.clh_memcpy_to_device.argtypes = [vp, vp, C.c_size_t]
    L.clh_memcpy_to_host.argtypes = [vp, vp, C.c_size_t]
    L.clh_cones_create.argtypes = [C.POINTER(vp), C.c_int32, ip, ip, dp, C.POINTER(ClhSettings)]
    L.clh_cones_get_info.argtypes = [vp, C.POINTER(ClhConesInfo)]
    L.clh_cones_update_scaling.argtypes = [vp, vp, vp, C.c_double, C.c_int32, vp, vp]
    L.clh_cones_set_identity_scaling.argtypes = [vp, vp]
    L.clh_cones_get_Hs.argtypes = [vp, vp, vp]
    L.clh_cones_mul_Hs.argtypes = [vp, vp, vp, vp]
    L.clh_cones_affine_ds.argtypes = [vp, vp, vp, vp]
    L.clh_cones_combined_ds_shift.argtypes = [vp, vp, vp, vp, C.c_double, vp]
    L.clh_cones_ds_from_dz_offset.argtypes = [vp, vp, vp, vp, vp]
    L.clh_cones_step_length.argtypes = [vp, vp, vp, vp, vp, C.c_double, vp, vp]
    L.clh_cones_step_length_noalloc.argtypes = [vp, vp, vp, vp, vp, C.c_double, vp, vp]
    L.clh_cones_compute_barrier.argtypes = [vp, vp, vp, vp, vp, C.c_double, vp, vp]
    L.clh_cones_margins.argtypes = [vp, vp, vp, vp]
    L.clh_cones_scaled_unit_shift.argtypes = [vp, vp, C.c_double, C.c_int32, vp]
    L.clh_cones_unit_initialization.argtypes = [vp, vp, vp, vp]
    L.clh_cones_rectify_equilibration.argtypes = [vp, vp, vp, ip, vp]
    L.clh_cones_destroy.argtypes = [vp]
    L.clh_cones_destroy.restype = None
    return L


_LIB = None


def lib():
    """The HIP library of this package (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(DEFAULT_LIB):
            raise RuntimeError(
                f"{DEFAULT_LIB} not found: build it with `python -m cuclarabel_amd.build` "
                "(hipcc, gfx950).  There is no CPU fallback.")
        _LIB = bind(DEFAULT_LIB)
    return _LIB


def check(L, rc):
    if rc != 0:
        msg = L.clh_last_error().decode()
        if rc < 0:
            raise ValueError(msg)
        raise RuntimeError(msg)
    return rc
