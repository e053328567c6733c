"""``compile_qdldl_variant(..., backend="hip")``: the HIP drop-in for the Warp backend.

Mirrors ``compile_qdldl_variant`` (``warpmpc/jax_qdldl/core.py:544-650``) and
``compile_qdldl_warp`` (``warpmpc/jax_qdldl/warp_backend.py:1175-1429``): same
keyword arguments, same ``CompiledQDLDL`` contract in the symbolic layout
(``values (nnz_a, B)``, ``rhs (n, B)``), same error behaviour for bad options.
Arrays are PyTorch-ROCm tensors on the current GPU; numpy input is copied there.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from .._device import as_device_f32, ptr, require_gpu, stream_ptr, torch
from .segments import make_level_schedule, make_segments, make_solve_levels
from .tables import ragged_program
from .types import CompiledQDLDL, QDLDLPlan

SUPPORTED_BACKENDS = ("hip",)


def normalize_backend_name(name: str, *, role: str) -> str:
    """Backend-name check (``core.py:513-516``); this build provides ``"hip"``."""
    if name not in SUPPORTED_BACKENDS:
        raise ValueError(
            f"unknown QDLDL {role} backend {name!r}; this build provides {SUPPORTED_BACKENDS}"
        )
    return name


class DevicePlan:
    """Owns the ``wmpc_qdldl_plan`` (device-resident tables) of one QDLDLPlan."""

    def __init__(self, plan: QDLDLPlan):
        lib = _lib.load()
        self.plan = plan
        self.tables = ragged_program(plan)
        pat = _lib.QdldlPattern()
        pat.n, pat.nnz_a, pat.nnz_l = plan.n, plan.nnz_a, plan.nnz_l
        for name, arr in self.tables.items():
            setattr(pat, name, _lib.i32p(arr))
        self.pattern = pat
        handle = ctypes.c_void_p()
        _lib.check(lib.wmpc_qdldl_plan_create(ctypes.byref(pat), ctypes.byref(handle)), "wmpc_qdldl_plan_create")
        self.handle = handle
        info = [ctypes.c_int32() for _ in range(4)]
        _lib.check(lib.wmpc_qdldl_plan_info(handle, *[ctypes.byref(v) for v in info]), "wmpc_qdldl_plan_info")
        self.max_reach = int(info[3].value)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().wmpc_qdldl_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


def _variant_name(*, transpose_work, segmented, segment_budget, segment_strategy,
                  level_scheduled_solve, level_scheduled_solve_threshold) -> str:
    parts = ["hip"]
    if transpose_work:
        parts.append("transpose")
    if segmented:
        parts += ["segmented", f"budget{segment_budget}", segment_strategy]
    if level_scheduled_solve:
        parts.append("levelsolve")
        if level_scheduled_solve_threshold != 1:
            parts.append(f"threshold{level_scheduled_solve_threshold}")
    return "+".join(parts)


def compile_qdldl_variant(
    plan: QDLDLPlan,
    dtype=np.float32,
    *,
    backend: str = "hip",
    factor_backend: str | None = None,
    solve_backend: str | None = None,
    transpose_work: bool = True,
    segmented: bool = False,
    segment_budget: int = 64,
    segment_strategy: str = "optimal",
    level_scheduled_solve: bool = False,
    level_scheduled_solve_threshold: int = 1,
) -> CompiledQDLDL:
    """Bind the HIP factor / solve kernels to ``plan``.

    Like the Warp backend, the HIP backend works in the symbolic layout and
    requires ``transpose_work=True`` and float32 (``warp_backend.py:1197-1228``).
    The segment and level options are validated and recorded in ``variant``; the
    HIP kernels walk unpadded row programs, so they need no padding segments.
    """
    dtype = np.dtype(dtype)
    factor_backend = normalize_backend_name(backend if factor_backend is None else factor_backend, role="factor")
    solve_backend = normalize_backend_name(backend if solve_backend is None else solve_backend, role="solve")
    if dtype != np.dtype(np.float32):
        raise ValueError("the HIP QDLDL backend currently supports float32")
    if not transpose_work:
        raise ValueError("the HIP QDLDL backend requires transpose_work=True")
    if level_scheduled_solve_threshold < 1:
        raise ValueError("level_scheduled_solve_threshold must be at least 1")
    if segmented:
        make_segments(plan.process_mask, segment_budget=segment_budget, segment_strategy=segment_strategy)
        make_segments(plan.col_mask, segment_budget=segment_budget, segment_strategy=segment_strategy)
    levels = make_solve_levels(plan.col_rows, plan.col_mask)
    schedule = make_level_schedule(levels, serial_width_threshold=level_scheduled_solve_threshold)
    require_gpu()
    dev = DevicePlan(plan)
    lib = _lib.load()
    th = torch()
    n, nnz_l = plan.n, plan.nnz_l

    def _check_rows(x, rows: int, name: str):
        if x.dim() != 2 or x.shape[0] != rows:
            raise ValueError(f"{name} must have shape ({rows}, batch), got {tuple(x.shape)}")

    def factor(a_values):
        a = as_device_f32(a_values, "a_values")
        _check_rows(a, plan.nnz_a, "a_values")
        b = int(a.shape[1])
        lx = th.empty((nnz_l, b), dtype=th.float32, device=a.device)
        d = th.empty((n, b), dtype=th.float32, device=a.device)
        dinv = th.empty((n, b), dtype=th.float32, device=a.device)
        _lib.check(lib.wmpc_qdldl_factor_f32(dev.handle, ptr(a), ptr(lx), ptr(d), ptr(dinv), b, b, stream_ptr()),
                   "wmpc_qdldl_factor_f32")
        return lx, d, dinv

    def solve(lx, dinv, rhs):
        lx = as_device_f32(lx, "lx")
        dinv = as_device_f32(dinv, "dinv")
        r = as_device_f32(rhs, "rhs")
        _check_rows(lx, nnz_l, "lx")
        _check_rows(dinv, n, "dinv")
        _check_rows(r, n, "rhs")
        b = int(r.shape[1])
        fb = int(lx.shape[1])
        if fb != dinv.shape[1] or fb not in (1, b):
            raise ValueError("lx/dinv batch must be 1 (shared) or equal to the rhs batch")
        f_ld = 0 if (fb == 1 and b != 1) else b
        x = th.empty((n, b), dtype=th.float32, device=r.device)
        work = th.empty((n, b), dtype=th.float32, device=r.device)
        _lib.check(lib.wmpc_qdldl_solve_f32(dev.handle, ptr(lx), f_ld, ptr(dinv), f_ld, ptr(r), ptr(x), ptr(work),
                                            work.numel(), b, b, stream_ptr()), "wmpc_qdldl_solve_f32")
        return x

    def factor_and_solve(a_values, rhs):
        a = as_device_f32(a_values, "a_values")
        r = as_device_f32(rhs, "rhs")
        _check_rows(a, plan.nnz_a, "a_values")
        _check_rows(r, n, "rhs")
        b = int(a.shape[1])
        if r.shape[1] != b:
            raise ValueError("a_values and rhs must have the same batch")
        x = th.empty((n, b), dtype=th.float32, device=a.device)
        lx = th.empty((nnz_l, b), dtype=th.float32, device=a.device)
        d = th.empty((n, b), dtype=th.float32, device=a.device)
        dinv = th.empty((n, b), dtype=th.float32, device=a.device)
        work = th.empty((n, b), dtype=th.float32, device=a.device)
        _lib.check(lib.wmpc_qdldl_factor_solve_f32(dev.handle, ptr(a), ptr(r), ptr(x), ptr(lx), ptr(d), ptr(dinv),
                                                   ptr(work), work.numel(), b, b, stream_ptr()),
                   "wmpc_qdldl_factor_solve_f32")
        return x, lx, d

    return CompiledQDLDL(
        plan=plan,
        dtype=dtype,
        factor=factor,
        solve=solve,
        factor_and_solve=factor_and_solve,
        variant=_variant_name(
            transpose_work=transpose_work,
            segmented=segmented,
            segment_budget=segment_budget,
            segment_strategy=segment_strategy,
            level_scheduled_solve=level_scheduled_solve,
            level_scheduled_solve_threshold=level_scheduled_solve_threshold,
        ),
        values_layout="original_symbolic",
        rhs_layout="symbolic",
        backend="hip",
        factor_backend=factor_backend,
        solve_backend=solve_backend,
        extras={"device_plan": dev, "solve_levels": levels, "level_schedule": schedule},
    )


def compile_qdldl(plan: QDLDLPlan, dtype=np.float32) -> CompiledQDLDL:
    """``compile_qdldl`` (``core.py:314-317``) on the HIP backend."""
    return compile_qdldl_variant(plan, dtype=dtype)


__all__ = ["DevicePlan", "compile_qdldl", "compile_qdldl_variant", "normalize_backend_name"]
