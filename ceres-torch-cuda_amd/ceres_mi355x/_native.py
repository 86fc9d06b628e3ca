"""ctypes binding of libceres_mi355x.so (the C ABI in include/ceres_mi355x.h).

The library is the only compute path: if it is missing or fails to load, every
solver entry point raises ``NativeLibraryError``; nothing falls back to a CPU
or PyTorch implementation.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(os.environ.get("CERES_MI355X_LIB", Path(__file__).resolve().parent / "libceres_mi355x.so"))

CT_OK = 0
CT_ERR_ARGUMENT = -1
CT_ERR_HIP = -2
CT_ERR_NOT_SPD = -3
CT_ERR_WORKSPACE = -4
CT_ERR_UNSUPPORTED = -5

CT_REC_FIELDS = 12
ABI_VERSION = 14  # CT_ABI_VERSION of include/ceres_mi355x.h
KERNELS = ["eval_obs", "eval_points", "camera_items", "camera_reduce", "lm_begin", "rcs_solve", "refine", "dogleg",
           "step", "candidate", "lm_end"]
REC = dict(
    iteration=0,
    cost=1,
    cost_change=2,
    gradient_max_norm=3,
    gradient_norm=4,
    step_norm=5,
    relative_decrease=6,
    trust_region_radius=7,
    step_is_valid=8,
    step_is_successful=9,
    linear_solver_iterations=10,
    timestamp=11,
)


# line-search iteration records (CT_LS_REC_*)
CT_LS_REC_FIELDS = 13
LS_REC = dict(
    iteration=0,
    cost=1,
    gradient_norm=2,
    gradient_max_norm=3,
    step_size=4,
    line_search_iterations=5,
    line_search_function_evaluations=6,
    line_search_gradient_evaluations=7,
    line_search_direction_restarts=8,
    step_norm=9,
    cost_change=10,
    step_is_successful=11,
    timestamp=12,
)


class NativeLibraryError(RuntimeError):
    """The HIP library is missing, failed to load, or a call failed."""


class ct_loss(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("a", ctypes.c_double), ("b", ctypes.c_double), ("scale", ctypes.c_double)]


class ct_lm_options(ctypes.Structure):
    _fields_ = [
        ("max_num_iterations", ctypes.c_int32),
        ("max_num_consecutive_invalid_steps", ctypes.c_int32),
        ("linear_solver", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("loop_mode", ctypes.c_int32),
        ("min_linear_solver_iterations", ctypes.c_int32),
        ("max_linear_solver_iterations", ctypes.c_int32),
        ("max_num_refinement_iterations", ctypes.c_int32),
        ("jacobi_scaling", ctypes.c_int32),
        ("iterations_per_graph", ctypes.c_int32),
        ("initial_trust_region_radius", ctypes.c_double),
        ("max_trust_region_radius", ctypes.c_double),
        ("min_trust_region_radius", ctypes.c_double),
        ("min_relative_decrease", ctypes.c_double),
        ("min_lm_diagonal", ctypes.c_double),
        ("max_lm_diagonal", ctypes.c_double),
        ("function_tolerance", ctypes.c_double),
        ("gradient_tolerance", ctypes.c_double),
        ("parameter_tolerance", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("trust_region_strategy", ctypes.c_int32),
        ("dogleg_type", ctypes.c_int32),
        ("preconditioner", ctypes.c_int32),
        ("max_solver_time_in_seconds", ctypes.c_double),
        ("debug_invalid_steps", ctypes.c_int32),
        ("schur_operator", ctypes.c_int32),
        ("debug_fill_lds", ctypes.c_int32),
        ("solver_workgroups", ctypes.c_int32),
    ]


class ct_lm_report(ctypes.Structure):
    _fields_ = [
        ("termination", ctypes.c_int32),
        ("reason", ctypes.c_int32),
        ("num_records", ctypes.c_int32),
        ("num_successful_steps", ctypes.c_int32),
        ("num_unsuccessful_steps", ctypes.c_int32),
        ("num_residual_evaluations", ctypes.c_int32),
        ("num_jacobian_evaluations", ctypes.c_int32),
        ("num_linear_solves", ctypes.c_int32),
        ("initial_cost", ctypes.c_double),
        ("final_cost", ctypes.c_double),
        ("device_seconds", ctypes.c_double),
    ]


class ct_line_search_options(ctypes.Structure):
    _fields_ = [
        ("max_num_iterations", ctypes.c_int32),
        ("direction", ctypes.c_int32),
        ("line_search_type", ctypes.c_int32),
        ("nonlinear_cg_type", ctypes.c_int32),
        ("max_lbfgs_rank", ctypes.c_int32),
        ("use_approximate_eigenvalue_bfgs_scaling", ctypes.c_int32),
        ("interpolation", ctypes.c_int32),
        ("max_num_step_size_iterations", ctypes.c_int32),
        ("max_num_direction_restarts", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("min_line_search_step_size", ctypes.c_double),
        ("sufficient_function_decrease", ctypes.c_double),
        ("max_step_contraction", ctypes.c_double),
        ("min_step_contraction", ctypes.c_double),
        ("sufficient_curvature_decrease", ctypes.c_double),
        ("function_tolerance", ctypes.c_double),
        ("gradient_tolerance", ctypes.c_double),
        ("parameter_tolerance", ctypes.c_double),
        ("max_solver_time_in_seconds", ctypes.c_double),
    ]


class ct_line_search_report(ctypes.Structure):
    _fields_ = [
        ("termination", ctypes.c_int32),
        ("reason", ctypes.c_int32),
        ("num_records", ctypes.c_int32),
        ("num_successful_steps", ctypes.c_int32),
        ("num_unsuccessful_steps", ctypes.c_int32),
        ("num_jacobian_evaluations", ctypes.c_int32),
        ("num_line_search_steps", ctypes.c_int32),
        ("num_line_search_function_evaluations", ctypes.c_int32),
        ("num_line_search_gradient_evaluations", ctypes.c_int32),
        ("num_line_search_direction_restarts", ctypes.c_int32),
        ("initial_cost", ctypes.c_double),
        ("final_cost", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]


_vp = ctypes.c_void_p


class ct_ba_problem(ctypes.Structure):
    _fields_ = [
        ("num_cameras", ctypes.c_int32),
        ("num_points", ctypes.c_int32),
        ("num_observations", ctypes.c_int32),
        ("num_blocks", ctypes.c_int32),
        ("num_terms", ctypes.c_int64),
        ("obs_camera", _vp),
        ("obs_point", _vp),
        ("point_offsets", _vp),
        ("observations", _vp),
        ("camera_offsets", _vp),
        ("camera_obs", _vp),
        ("block_cameras", _vp),
        ("block_offsets", _vp),
        ("block_terms", _vp),
        ("camera_mask", _vp),
        ("point_mask", _vp),
        ("loss", ct_loss),
        ("camera_model", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ct_curve_batch(ctypes.Structure):
    _fields_ = [
        ("num_problems", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("num_samples", ctypes.c_int64),
        ("offsets", _vp),
        ("x", _vp),
        ("y", _vp),
        ("order", _vp),
        ("loss", ct_loss),
        ("x32", _vp),
        ("y32", _vp),
    ]


class ct_pose2d_problem(ctypes.Structure):
    _fields_ = [
        ("num_nodes", ctypes.c_int32),
        ("num_edges", ctypes.c_int32),
        ("num_pairs", ctypes.c_int32),
        ("fixed_node", ctypes.c_int32),
        ("edges", _vp),
        ("measurements", _vp),
        ("node_offsets", _vp),
        ("node_edges", _vp),
        ("pair_nodes", _vp),
        ("pair_offsets", _vp),
        ("pair_edges", _vp),
        ("loss", ct_loss),
    ]


class ct_ba_batch_problem(ctypes.Structure):
    _fields_ = [
        ("num_cameras", ctypes.c_int32),
        ("num_points", ctypes.c_int32),
        ("num_observations", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("num_terms", ctypes.c_int64),
        ("camera_base", ctypes.c_int64),
        ("point_base", ctypes.c_int64),
        ("obs_base", ctypes.c_int64),
        ("point_offsets_base", ctypes.c_int64),
        ("camera_offsets_base", ctypes.c_int64),
        ("block_offsets_base", ctypes.c_int64),
        ("term_base", ctypes.c_int64),
        ("work_base", ctypes.c_int64),
        ("loss", ct_loss),
    ]


class ct_ba_batch(ctypes.Structure):
    _fields_ = [
        ("num_problems", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("problems", _vp),
        ("obs_camera", _vp),
        ("obs_point", _vp),
        ("observations", _vp),
        ("point_offsets", _vp),
        ("camera_offsets", _vp),
        ("camera_obs", _vp),
        ("block_offsets", _vp),
        ("block_terms", _vp),
        ("order", _vp),
    ]


CT_BA_BATCH_MAX_CAMERAS = 21


class ct_batch_schedule(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int32), ("groups", ctypes.c_int32), ("wave_problems", ctypes.c_int32)]


# numpy view of a device array of ct_lm_report
import numpy as _np  # noqa: E402

REPORT_DTYPE = _np.dtype([(name, _np.int32 if ctype is ctypes.c_int32 else _np.float64)
                          for name, ctype in ct_lm_report._fields_], align=True)
assert REPORT_DTYPE.itemsize == ctypes.sizeof(ct_lm_report)

# (name, restype, argtypes) for every entry point the header declares
SIGNATURES = {
    "ct_abi_version": (ctypes.c_int, []),
    "ct_last_error": (ctypes.c_char_p, []),
    "ct_ba_workspace_bytes": (ctypes.c_int, [ctypes.POINTER(ct_ba_problem), ctypes.POINTER(ct_lm_options),
                                             ctypes.POINTER(ctypes.c_size_t)]),
    "ct_ba_plan_create": (ctypes.c_int, [ctypes.POINTER(ct_ba_problem), ctypes.POINTER(ct_lm_options), _vp, _vp, _vp,
                                         ctypes.c_size_t, _vp, ctypes.POINTER(_vp)]),
    "ct_ba_plan_run": (ctypes.c_int, [_vp, ctypes.POINTER(ct_lm_report), _vp]),
    "ct_ba_plan_destroy": (ctypes.c_int, [_vp]),
    "ct_ba_plan_profile": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "ct_ba_evaluate": (ctypes.c_int, [ctypes.POINTER(ct_ba_problem), _vp, _vp, ctypes.POINTER(ctypes.c_double), _vp,
                                      _vp, ctypes.c_size_t, _vp]),
    "ct_ba_evaluate_jacobian": (ctypes.c_int, [ctypes.POINTER(ct_ba_problem), _vp, _vp, _vp, _vp, _vp,
                                               ctypes.c_int64, ctypes.c_int32, _vp, _vp,
                                               ctypes.POINTER(ctypes.c_double), _vp, ctypes.c_size_t, _vp]),
    "ct_curve_workspace_bytes": (ctypes.c_int, [ctypes.POINTER(ct_curve_batch), ctypes.POINTER(ctypes.c_size_t)]),
    "ct_curve_solve": (ctypes.c_int, [ctypes.POINTER(ct_curve_batch), ctypes.POINTER(ct_lm_options),
                                      ctypes.POINTER(ct_batch_schedule), _vp, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "ct_curve_validate": (ctypes.c_int, [ctypes.POINTER(ct_curve_batch), ctypes.POINTER(ctypes.c_int32)]),
    "ct_ba_batch_problem_bytes": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int64,
                                                 ctypes.POINTER(ctypes.c_size_t)]),
    "ct_ba_batch_tables": (ctypes.c_int, [ctypes.c_int32] + [_vp] * 14),
    "ct_ba_batch_solve": (ctypes.c_int, [ctypes.POINTER(ct_ba_batch), ctypes.POINTER(ct_lm_options), _vp, _vp, _vp,
                                         _vp, _vp, ctypes.c_size_t, _vp]),
    "ct_pose2d_workspace_bytes": (ctypes.c_int, [ctypes.POINTER(ct_pose2d_problem), ctypes.POINTER(ct_lm_options),
                                                 ctypes.POINTER(ctypes.c_size_t)]),
    "ct_pose2d_plan_create": (ctypes.c_int, [ctypes.POINTER(ct_pose2d_problem), ctypes.POINTER(ct_lm_options), _vp,
                                             _vp, ctypes.c_size_t, _vp, ctypes.POINTER(_vp)]),
    "ct_pose2d_plan_run": (ctypes.c_int, [_vp, ctypes.POINTER(ct_lm_report), _vp]),
    "ct_pose2d_plan_system_size": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "ct_pose2d_plan_destroy": (ctypes.c_int, [_vp]),
    "ct_ba_covariance_workspace_bytes": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    "ct_ba_plan_covariance": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_double, _vp,
                                             ctypes.POINTER(ctypes.c_int64), _vp, ctypes.c_size_t]),
    "ct_ba_line_search_workspace_bytes": (ctypes.c_int, [ctypes.POINTER(ct_ba_problem),
                                                         ctypes.POINTER(ct_line_search_options),
                                                         ctypes.POINTER(ctypes.c_size_t)]),
    "ct_ba_line_search_solve": (ctypes.c_int, [ctypes.POINTER(ct_ba_problem), ctypes.POINTER(ct_line_search_options),
                                               _vp, _vp, _vp, ctypes.c_size_t, _vp,
                                               ctypes.POINTER(ct_line_search_report), _vp]),
    "ct_dense_workspace_bytes": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    "ct_normal_equations_solve_f64": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _vp, _vp,
                                                     _vp, _vp, ctypes.c_size_t, _vp]),
    "ct_normal_equations_solve_f64_noalloc": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                             _vp, _vp, _vp, _vp]),
    "ct_block_schur_solve_f64": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _vp,
                                                ctypes.c_int64, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
    "ct_rcs_workspace_bytes": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    "ct_rcs_potrf_f64": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, ctypes.c_size_t, _vp]),
    "ct_rcs_potrs_f64": (ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_size_t, _vp]),
}

_LIB = None
_LOAD_ERROR: str | None = None


def load() -> ctypes.CDLL:
    """Load (once) and return the native library, or raise NativeLibraryError."""
    global _LIB, _LOAD_ERROR
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        _LOAD_ERROR = f"{LIB_PATH} not found; build it with ceres-torch-cuda_amd/build.py"
        raise NativeLibraryError(_LOAD_ERROR)
    try:
        lib = ctypes.CDLL(str(LIB_PATH))
    except OSError as exc:  # pragma: no cover - depends on the host
        _LOAD_ERROR = f"failed to load {LIB_PATH}: {exc}"
        raise NativeLibraryError(_LOAD_ERROR) from exc
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    if lib.ct_abi_version() != ABI_VERSION:
        raise NativeLibraryError("ABI version mismatch")
    _LIB = lib
    return lib


def check(status: int, what: str) -> None:
    if status != CT_OK:
        msg = load().ct_last_error()
        raise NativeLibraryError(f"{what} failed ({status}): {msg.decode() if msg else ''}")


def exported_symbols() -> list[str]:
    return list(SIGNATURES)


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    return int(t.data_ptr())


def current_stream_handle() -> int | None:
    import torch

    return int(torch.cuda.current_stream().cuda_stream) or None
