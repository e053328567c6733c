"""``compile_qdldl_variant(..., backend="hip")``: the HIP drop-in for the Warp backend.

Mirrors ``compile_qdldl_variant`` (``warpmpc/jax_qdldl/core.py:544-650``) and
``compile_qdldl_warp`` (``warpmpc/jax_qdldl/warp_backend.py:1175-1429``): same
keyword arguments, same ``CompiledQDLDL`` contract in the symbolic layout
(``values (nnz_a, B)``, ``rhs (n, B)``), same error behaviour for bad options.
Arrays are PyTorch-ROCm tensors on the current GPU; numpy input is copied there.
"""

from __future__ import annotations

import ctypes
import warnings

import numpy as np

from .. import _lib
from .._device import as_device_f32, ptr, require_gpu, stream_ptr, torch
from .segments import make_segments
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
        self.ipw_setting = 0

    def set_instances_per_wave(self, ipw: int) -> None:
        """Instances per wavefront of this plan's launches (0: the process-wide choice)."""
        _lib.check(_lib.load().wmpc_qdldl_plan_set_instances_per_wave(self.handle, int(ipw)),
                   "wmpc_qdldl_plan_set_instances_per_wave")
        self.ipw_setting = int(ipw)

    def set_lds_budget(self, nbytes: int) -> None:
        """LDS bytes per workgroup of this plan's kernels (re-derives its programs)."""
        _lib.check(_lib.load().wmpc_qdldl_plan_set_lds_budget(self.handle, int(nbytes)),
                   "wmpc_qdldl_plan_set_lds_budget")

    @property
    def lds_budget(self) -> int:
        return int(_lib.load().wmpc_qdldl_plan_lds_budget(self.handle))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().wmpc_qdldl_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


# The kernels that run: unpadded per-row programs, one lane per instance,
# batch-major ("transpose") work arrays.
VARIANT = "hip+transpose"

_warned_options: set[str] = set()


def _note_options_without_effect(segmented: bool) -> dict:
    """``segmented`` chooses how the Warp kernels pad their per-thread loops
    (``warp_backend.py:1129-1149``).  It never changes results, and the HIP
    kernels walk unpadded programs, so there is nothing for it to select: it
    is validated, recorded as without effect and reported once per process."""
    if segmented and "segmented" not in _warned_options:
        _warned_options.add("segmented")
        warnings.warn("segmented=True has no effect on the HIP QDLDL backend (its kernels have no padded loops)",
                      UserWarning, stacklevel=3)
    return {"segmented": True} if segmented else {}


def level_lanes(n: int, batch: int) -> int:
    """Lanes per instance of the level-scheduled solve: 8, or more when the
    LDS of 64 / lanes instance vectors of n floats would exceed 64 KiB."""
    lanes = 8
    while lanes < 64 and n * (64 // lanes) * 4 > 64 * 1024:
        lanes *= 2
    return lanes


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
    requires ``transpose_work=True`` (``warp_backend.py:1197-1228``); it
    computes in float32 or float64.  ``level_scheduled_solve=True`` makes
    ``solve`` (and the solve of ``factor_and_solve``) the level-scheduled
    kernel: ``level_lanes`` lanes per instance, levels wider than
    ``level_scheduled_solve_threshold`` shared by them, narrower runs serial
    (float32).  ``segmented`` is validated like the reference validates it and
    has no effect (``_note_options_without_effect``); ``variant`` names the
    kernels that run.
    """
    dtype = np.dtype(dtype)
    factor_backend = normalize_backend_name(backend if factor_backend is None else factor_backend, role="factor")
    solve_backend = normalize_backend_name(backend if solve_backend is None else solve_backend, role="solve")
    if dtype not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ValueError(f"the HIP QDLDL backend supports float32 and float64, not {dtype}")
    f64 = dtype == np.dtype(np.float64)
    if not transpose_work:
        raise ValueError("the HIP QDLDL backend requires transpose_work=True")
    if level_scheduled_solve_threshold < 1:
        raise ValueError("level_scheduled_solve_threshold must be at least 1")
    if level_scheduled_solve and f64:
        raise ValueError("the level-scheduled HIP solve computes in float32")
    if segmented:
        make_segments(plan.process_mask, segment_budget=segment_budget, segment_strategy=segment_strategy)
        make_segments(plan.col_mask, segment_budget=segment_budget, segment_strategy=segment_strategy)
    without_effect = _note_options_without_effect(segmented)
    require_gpu()
    dev = DevicePlan(plan)
    lib = _lib.load()
    th = torch()
    n, nnz_l = plan.n, plan.nnz_l
    tdt = th.float64 if f64 else th.float32

    def _dev(x, name):
        if f64:
            t = x if isinstance(x, th.Tensor) else th.as_tensor(np.asarray(x, dtype=np.float64))
            return t.to(device="cuda", dtype=th.float64).contiguous()
        return as_device_f32(x, name)

    def _check_rows(x, rows: int, name: str):
        if x.dim() != 2 or x.shape[0] != rows:
            raise ValueError(f"{name} must have shape ({rows}, batch), got {tuple(x.shape)}")

    def factor(a_values):
        a = _dev(a_values, "a_values")
        _check_rows(a, plan.nnz_a, "a_values")
        b = int(a.shape[1])
        lx = th.empty((nnz_l, b), dtype=tdt, device=a.device)
        d = th.empty((n, b), dtype=tdt, device=a.device)
        dinv = th.empty((n, b), dtype=tdt, device=a.device)
        if f64:
            work = th.empty((max(n * b, 1),), dtype=tdt, device=a.device)
            _lib.check(lib.wmpc_qdldl_factor_f64(dev.handle, ptr(a), ptr(lx), ptr(d), ptr(dinv), ptr(work),
                                                 work.numel(), b, b, stream_ptr()), "wmpc_qdldl_factor_f64")
        else:
            _lib.check(lib.wmpc_qdldl_factor_f32(dev.handle, ptr(a), ptr(lx), ptr(d), ptr(dinv), b, b,
                                                 stream_ptr()), "wmpc_qdldl_factor_f32")
        return lx, d, dinv

    def solve(lx, dinv, rhs):
        lx = _dev(lx, "lx")
        dinv = _dev(dinv, "dinv")
        r = _dev(rhs, "rhs")
        _check_rows(lx, nnz_l, "lx")
        _check_rows(dinv, n, "dinv")
        _check_rows(r, n, "rhs")
        b = int(r.shape[1])
        fb = int(lx.shape[1])
        if fb != dinv.shape[1] or fb not in (1, b):
            raise ValueError("lx/dinv batch must be 1 (shared) or equal to the rhs batch")
        f_ld = 0 if (fb == 1 and b != 1) else b
        x = th.empty((n, b), dtype=tdt, device=r.device)
        if level_scheduled_solve:
            _lib.check(lib.wmpc_qdldl_solve_levels_f32(dev.handle, int(level_scheduled_solve_threshold),
                                                       level_lanes(n, b), ptr(lx), f_ld, ptr(dinv), f_ld, ptr(r),
                                                       ptr(x), b, b, stream_ptr()), "wmpc_qdldl_solve_levels_f32")
            return x
        work = th.empty((n, b), dtype=tdt, device=r.device)
        fn = lib.wmpc_qdldl_solve_f64 if f64 else lib.wmpc_qdldl_solve_f32
        _lib.check(fn(dev.handle, ptr(lx), f_ld, ptr(dinv), f_ld, ptr(r), ptr(x), ptr(work), work.numel(), b, b,
                      stream_ptr()), "wmpc_qdldl_solve_f64" if f64 else "wmpc_qdldl_solve_f32")
        return x

    def factor_and_solve(a_values, rhs):
        if f64 or level_scheduled_solve:
            # two launches: the factorization, then the solve kernel chosen above
            lx, d, dinv = factor(a_values)
            r = _dev(rhs, "rhs")
            _check_rows(r, n, "rhs")
            if r.shape[1] != lx.shape[1]:
                raise ValueError("a_values and rhs must have the same batch")
            return solve(lx, dinv, r), lx, d
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
        variant=VARIANT + ("+f64" if f64 else "") + ("+levelsolve" if level_scheduled_solve else ""),
        values_layout="original_symbolic",
        rhs_layout="symbolic",
        backend="hip",
        factor_backend=factor_backend,
        solve_backend=solve_backend,
        extras={"device_plan": dev, "options_without_effect": without_effect},
    )


def compile_qdldl(plan: QDLDLPlan, dtype=np.float32) -> CompiledQDLDL:
    """``compile_qdldl`` (``core.py:314-317``) on the HIP backend."""
    return compile_qdldl_variant(plan, dtype=dtype)


__all__ = ["DevicePlan", "compile_qdldl", "compile_qdldl_variant", "normalize_backend_name"]
