"""ctypes binding of the HIP library ``libwarpmpc_hip.so`` (C ABI: include/warpmpc_hip.h).

The library is the only compute path: if it is missing or cannot be loaded,
every entry point raises instead of falling back to anything else.
Device memory and streams come from PyTorch-ROCm tensors (``data_ptr()`` and
``torch.cuda.current_stream().cuda_stream``); PyTorch is plumbing here.
"""

from __future__ import annotations

import ctypes
import pathlib
import threading

import numpy as np

_PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = _PKG / "libwarpmpc_hip.so"

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32

WMPC_OK = 0


class QdldlPattern(ctypes.Structure):
    """Mirror of ``wmpc_qdldl_pattern``."""

    _fields_ = [
        ("n", ctypes.c_int32),
        ("nnz_a", ctypes.c_int32),
        ("nnz_l", ctypes.c_int32),
        ("perm_to_orig", _i32p),
        ("diag_pos", _i32p),
        ("init_ptr", _i32p),
        ("init_rows", _i32p),
        ("init_pos", _i32p),
        ("proc_ptr", _i32p),
        ("proc_cols", _i32p),
        ("proc_lidx", _i32p),
        ("proc_off", _i32p),
        ("lp", _i32p),
        ("li", _i32p),
        ("p", _i32p),
        ("pinv", _i32p),
    ]


class OsqpDesc(ctypes.Structure):
    """Mirror of ``wmpc_osqp_desc``."""

    _fields_ = [
        ("n", ctypes.c_int32),
        ("m", ctypes.c_int32),
        ("nnz_p", ctypes.c_int32),
        ("nnz_a", ctypes.c_int32),
        ("nnz_kkt", ctypes.c_int32),
        ("nnz_p_full", ctypes.c_int32),
        ("sigma", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("kkt_base", _f64p),
        ("p_to_kkt", _i32p),
        ("p_diag_sigma", _f64p),
        ("a_to_kkt", _i32p),
        ("p_scale", _f64p),
        ("a_scale", _f64p),
        ("q_scale", _f64p),
        ("scaling_d", _f64p),
        ("scaling_dinv", _f64p),
        ("scaling_e", _f64p),
        ("scaling_einv", _f64p),
        ("y_scale", _f64p),
        ("y_unscale", _f64p),
        ("rho_vec", _f64p),
        ("rho_inv_vec", _f64p),
        ("a_rows", _i32p),
        ("a_cols", _i32p),
        ("p_full_rows", _i32p),
        ("p_full_cols", _i32p),
        ("p_full_value_map", _i32p),
    ]


class AdjointDesc(ctypes.Structure):
    """Mirror of ``wmpc_adjoint_desc``."""

    _fields_ = [
        ("n", ctypes.c_int32),
        ("m", ctypes.c_int32),
        ("n_l", ctypes.c_int32),
        ("n_u", ctypes.c_int32),
        ("n_eq", ctypes.c_int32),
        ("nnz_adj", ctypes.c_int32),
        ("nnz_p", ctypes.c_int32),
        ("nnz_a", ctypes.c_int32),
        ("n_p_full", ctypes.c_int32),
        ("n_g", ctypes.c_int32),
        ("n_aeq", ctypes.c_int32),
        ("eps", ctypes.c_double),
        ("adj_indptr", _i32p),
        ("adj_indices", _i32p),
        ("adj_data", _f64p),
        ("lower_indices", _i32p),
        ("upper_indices", _i32p),
        ("eq_indices", _i32p),
        ("p_upper_rows", _i32p),
        ("p_upper_cols", _i32p),
        ("p_full_value_map", _i32p),
        ("g_rows", _i32p),
        ("g_cols", _i32p),
        ("g_a_value_map", _i32p),
        ("g_sign", _f64p),
        ("aeq_a_value_map", _i32p),
        ("diag_first_pos", _i32p),
        ("diag_second_pos", _i32p),
        ("p_block_pos", _i32p),
        ("g_block_pos", _i32p),
        ("aeq_block_pos", _i32p),
        ("glambda_block_pos", _i32p),
        ("slack_diag_pos", _i32p),
        ("aeq_t_block_pos", _i32p),
        ("a_rows", _i32p),
        ("a_cols", _i32p),
    ]


class AdjointKktDesc(ctypes.Structure):
    """Mirror of ``wmpc_adjoint_kkt_desc``."""

    _fields_ = [
        ("n", ctypes.c_int32),
        ("m", ctypes.c_int32),
        ("nnz_p", ctypes.c_int32),
        ("nnz_a", ctypes.c_int32),
        ("nnz_kkt", ctypes.c_int32),
        ("n_l", ctypes.c_int32),
        ("n_u", ctypes.c_int32),
        ("n_eq", ctypes.c_int32),
        ("delta", ctypes.c_double),
        ("kkt_rows", _i32p),
        ("kkt_cols", _i32p),
        ("p_to_kkt", _i32p),
        ("p_diag_sigma", _f64p),
        ("a_to_kkt", _i32p),
        ("lower_indices", _i32p),
        ("upper_indices", _i32p),
        ("eq_indices", _i32p),
        ("p_upper_rows", _i32p),
        ("p_upper_cols", _i32p),
        ("a_rows", _i32p),
        ("a_cols", _i32p),
    ]


class HeteroIO(ctypes.Structure):
    """Mirror of ``wmpc_osqp_hetero_io``."""

    _fields_ = [
        ("p_values", _vp), ("a_values", _vp), ("pm_ld", ctypes.c_int64),
        ("q", _vp), ("l", _vp), ("u", _vp), ("x0", _vp), ("z0", _vp), ("y0", _vp),
        ("x", _vp), ("z", _vp), ("y", _vp), ("prim", _vp), ("dual", _vp), ("obj", _vp),
        ("batch", ctypes.c_int64), ("ld", ctypes.c_int64),
    ]


# name -> (restype, argtypes); every symbol include/warpmpc_hip.h declares
SIGNATURES: dict[str, tuple] = {
    "wmpc_last_error": (ctypes.c_char_p, []),
    "wmpc_version": (ctypes.c_char_p, []),
    "wmpc_qdldl_plan_create": (ctypes.c_int, [ctypes.POINTER(QdldlPattern), ctypes.POINTER(_vp)]),
    "wmpc_qdldl_plan_create_host": (ctypes.c_int, [ctypes.POINTER(QdldlPattern), ctypes.POINTER(_vp)]),
    "wmpc_qdldl_plan_destroy": (ctypes.c_int, [_vp]),
    "wmpc_qdldl_plan_set_lds_budget": (ctypes.c_int, [_vp, _i64]),
    "wmpc_qdldl_plan_info": (ctypes.c_int, [_vp, _i32p, _i32p, _i32p, _i32p]),
    "wmpc_qdldl_plan_slots": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), _i32]),
    "wmpc_qdldl_plan_program": (ctypes.c_int, [_vp, _i32, _i32p, _i64, ctypes.POINTER(ctypes.c_int64)]),
    "wmpc_set_instances_per_wave": (ctypes.c_int, [_i32]),
    "wmpc_instances_per_wave": (ctypes.c_int32, [_i64]),
    "wmpc_qdldl_factor_f32": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp]),
    "wmpc_qdldl_solve_f32": (ctypes.c_int, [_vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _i64, _i64, _i64, _vp]),
    "wmpc_qdldl_factor_solve_f32": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64,
                                                   _vp]),
    "wmpc_osqp_plan_create": (ctypes.c_int, [ctypes.POINTER(OsqpDesc), _vp, ctypes.POINTER(_vp)]),
    "wmpc_osqp_plan_destroy": (ctypes.c_int, [_vp]),
    "wmpc_osqp_pattern_key": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "wmpc_osqp_plan_specialize": (ctypes.c_int, [_vp, _vp, _i64]),
    "wmpc_osqp_plan_specialized": (ctypes.c_int, [_vp]),
    "wmpc_osqp_factor_f32": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp, _i64, _i64, _vp]),
    "wmpc_osqp_work_floats": (ctypes.c_int64, [_vp, _i32, _i64]),
    "wmpc_osqp_solve_f32": (ctypes.c_int, [
        _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i64,
        _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i64, _vp]),
    "wmpc_osqp_hetero_plan_create": (ctypes.c_int, [ctypes.POINTER(_vp), _i32, ctypes.POINTER(_vp)]),
    "wmpc_osqp_hetero_plan_destroy": (ctypes.c_int, [_vp]),
    "wmpc_osqp_hetero_work_floats": (ctypes.c_int64, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    "wmpc_osqp_solve_hetero_f32": (ctypes.c_int, [_vp, ctypes.POINTER(HeteroIO), _i32, _vp, _i64, _vp]),
    "wmpc_gather_rows_f32": (ctypes.c_int, [_vp, _i32, _i32, _vp, _i64, _vp, _i64, _i64, _vp]),
    "wmpc_osqp_noalloc_max_floats": (ctypes.c_int64, []),
    "wmpc_osqp_solve_f32_noalloc": (ctypes.c_int, [
        _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
        _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _i64, _vp]),
    "wmpc_qdldl_plan_set_instances_per_wave": (ctypes.c_int, [_vp, _i32]),
    "wmpc_qdldl_plan_instances_per_wave": (ctypes.c_int32, [_vp, _i64]),
    "wmpc_qdldl_plan_lds_budget": (ctypes.c_int64, [_vp]),
    "wmpc_qdldl_f64_work_doubles": (ctypes.c_int64, [_vp, _i64]),
    "wmpc_qdldl_factor_f64": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp]),
    "wmpc_qdldl_solve_f64": (ctypes.c_int, [_vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _i64, _i64, _i64, _vp]),
    "wmpc_qdldl_plan_levels": (ctypes.c_int32, [_vp]),
    "wmpc_qdldl_solve_levels_f32": (ctypes.c_int, [_vp, _i32, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64,
                                                   _vp]),
    "wmpc_adjoint_plan_create": (ctypes.c_int, [ctypes.POINTER(AdjointDesc), _vp, ctypes.POINTER(_vp)]),
    "wmpc_adjoint_plan_destroy": (ctypes.c_int, [_vp]),
    "wmpc_adjoint_kkt_plan_create": (ctypes.c_int, [ctypes.POINTER(AdjointKktDesc), _vp, ctypes.POINTER(_vp)]),
    "wmpc_adjoint_is_kkt_form": (ctypes.c_int32, [_vp]),
    "wmpc_adjoint_dim": (ctypes.c_int32, [_vp]),
    "wmpc_adjoint_work_floats": (ctypes.c_int64, [_vp, _i64]),
    "wmpc_osqp_adjoint_f32": (ctypes.c_int, [
        _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _i64,
        _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp]),
}

PROGRAM_TABLES = ("row", "init", "proc", "upd", "bcol", "bent", "fcol", "bcol2", "fent", "ftpos", "frec",
                  "brec")


class HipError(RuntimeError):
    """A C-ABI call returned an error code."""


def load(path: pathlib.Path | str | None = None) -> ctypes.CDLL:
    """Load (once) and type the HIP library; raises if it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        p = pathlib.Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise ImportError(
                f"{p} is missing: build the HIP library first (python -m warpmpc_amd.build)"
            )
        lib = ctypes.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def check(rc: int, what: str) -> None:
    if rc != WMPC_OK:
        msg = load().wmpc_last_error().decode(errors="replace")
        raise HipError(f"{what} failed (code {rc}): {msg}")


def plan_slots(handle) -> dict[str, int]:
    """LDS slot counts of a QDLDL plan (``wmpc_qdldl_plan_slots``)."""
    out = (ctypes.c_int64 * 10)()
    check(load().wmpc_qdldl_plan_slots(handle, out, 10), "wmpc_qdldl_plan_slots")
    keys = ("reach", "colwin", "live_l", "backwin", "live_l_needed", "global_rereads", "lds_factor_solve",
            "lds_solve", "lds_budget", "lds_columns")
    return dict(zip(keys, (int(v) for v in out)))


def plan_program(handle, name: str) -> np.ndarray:
    """Host copy of one program table of a QDLDL plan (``wmpc_qdldl_plan_program``)."""
    which = PROGRAM_TABLES.index(name)
    n = ctypes.c_int64()
    lib = load()
    check(lib.wmpc_qdldl_plan_program(handle, which, None, 0, ctypes.byref(n)), "wmpc_qdldl_plan_program")
    out = np.zeros(max(int(n.value), 1), dtype=np.int32)
    check(lib.wmpc_qdldl_plan_program(handle, which, out.ctypes.data_as(_i32p), out.size, ctypes.byref(n)),
          "wmpc_qdldl_plan_program")
    return out[: int(n.value)]


def i32p(a: np.ndarray):
    return a.ctypes.data_as(_i32p)


def f64p(a: np.ndarray):
    return a.ctypes.data_as(_f64p)


__all__ = ["AdjointDesc", "AdjointKktDesc", "HeteroIO", "HipError", "LIB_PATH", "OsqpDesc", "PROGRAM_TABLES", "QdldlPattern", "SIGNATURES", "check",
           "load", "plan_program", "plan_slots"]
