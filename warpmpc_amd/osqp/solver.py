"""``compile_osqp(..., qdldl_backend="hip")``: fixed-iteration batched OSQP on MI355X.

Mirrors ``compile_osqp`` (``warpmpc/jax_osqp/solver.py:685-1492``): same keyword
arguments and the same ``CompiledOSQP`` callables with batch-first arrays
(``p_values (B|1, nnz_p)``, ``a_values (B|1, nnz_a)``, ``q (B, n)``, ``l, u (B, m)``).
Where the reference runs ``lax.scan`` over ``_iterate`` with Warp solves in
between, the HIP backend runs the whole solve — KKT assembly, factorization,
``max_iter`` ADMM iterations, unscaling, residuals — in one fused kernel launch
(``wmpc_osqp_solve_f32``).  Arrays are PyTorch-ROCm tensors; numpy is copied in.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from .._device import as_device_f32, ptr, require_gpu, stream_ptr, torch
from ..qdldl.core import compile_qdldl_variant
from .types import CompiledOSQP, FixedOSQPPlan, OSQPWarmStart


class DeviceOSQPPlan:
    """Owns the ``wmpc_osqp_plan`` of one FixedOSQPPlan (references the KKT plan)."""

    def __init__(self, plan: FixedOSQPPlan, kkt_device_plan):
        lib = _lib.load()
        self.kkt = kkt_device_plan
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        self.keep = {
            "kkt_base": f64(plan.kkt_upper.data),
            "p_to_kkt": i32(plan.p_to_kkt),
            "p_diag_sigma": f64(plan.p_diag_sigma),
            "a_to_kkt": i32(plan.a_to_kkt),
            "p_scale": f64(plan.p_scale),
            "a_scale": f64(plan.a_scale),
            "q_scale": f64(plan.q_scale),
            "scaling_d": f64(plan.scaling_d),
            "scaling_dinv": f64(plan.scaling_dinv),
            "scaling_e": f64(plan.scaling_e),
            "scaling_einv": f64(plan.scaling_einv),
            "y_scale": f64(plan.y_scale),
            "y_unscale": f64(plan.y_unscale),
            "rho_vec": f64(plan.rho_vec),
            "rho_inv_vec": f64(plan.rho_inv_vec),
            "a_rows": i32(plan.a_rows),
            "a_cols": i32(plan.a_cols),
            "p_full_rows": i32(plan.p_full_rows),
            "p_full_cols": i32(plan.p_full_cols),
            "p_full_value_map": i32(plan.p_full_value_map),
        }
        d = _lib.OsqpDesc()
        d.n, d.m = plan.n, plan.m
        d.nnz_p, d.nnz_a, d.nnz_kkt = plan.nnz_p, plan.nnz_a, plan.nnz_kkt
        d.nnz_p_full = int(plan.p_full_rows.size)
        d.sigma = float(plan.settings.sigma)
        d.alpha = float(plan.settings.alpha)
        for name, arr in self.keep.items():
            setattr(d, name, _lib.f64p(arr) if arr.dtype == np.float64 else _lib.i32p(arr))
        self.desc = d
        handle = ctypes.c_void_p()
        _lib.check(lib.wmpc_osqp_plan_create(ctypes.byref(d), kkt_device_plan.handle, ctypes.byref(handle)),
                   "wmpc_osqp_plan_create")
        self.handle = handle

    def pattern_key(self) -> int:
        key = ctypes.c_uint64()
        _lib.check(_lib.load().wmpc_osqp_pattern_key(self.handle, ctypes.byref(key)), "wmpc_osqp_pattern_key")
        return int(key.value)

    def specialize(self, code_object: bytes | None) -> None:
        """Attaches a generated straight-line kernel (code object bytes), or
        detaches with None."""
        lib = _lib.load()
        if code_object is None:
            _lib.check(lib.wmpc_osqp_plan_specialize(self.handle, None, 0), "wmpc_osqp_plan_specialize")
            return
        buf = ctypes.create_string_buffer(code_object, len(code_object))
        _lib.check(lib.wmpc_osqp_plan_specialize(self.handle, ctypes.cast(buf, ctypes.c_void_p), len(code_object)),
                   "wmpc_osqp_plan_specialize")

    @property
    def specialized(self) -> bool:
        r = int(_lib.load().wmpc_osqp_plan_specialized(self.handle))
        if r < 0:
            _lib.check(r, "wmpc_osqp_plan_specialized")
        return r == 1

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().wmpc_osqp_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


def _batched(x, width: int, name: str):
    x = as_device_f32(x, name)
    if x.dim() == 1:
        x = x[None, :]
    if x.dim() != 2 or x.shape[1] != width:
        raise ValueError(f"{name} must have shape (batch, {width}), got {tuple(x.shape)}")
    return x


def _symbolic(x, batch: int):
    """(1|B, w) batch-first -> contiguous (w, B) batch-major."""
    if x.shape[0] != batch:
        x = x.expand(batch, x.shape[1])
    return x.t().contiguous()


def compile_osqp(
    plan: FixedOSQPPlan,
    dtype=np.float32,
    *,
    qdldl_backend: str = "hip",
    qdldl_factor_backend: str | None = None,
    qdldl_solve_backend: str | None = None,
    transpose_work: bool = True,
    segmented: bool = False,
    segment_budget: int = 64,
    segment_strategy: str = "optimal",
    level_scheduled_solve: bool = False,
    level_scheduled_solve_threshold: int = 1,
    derivatives: bool = False,
    derivative_refinement_iters: int = 100,
    specialize: bool | str = "auto",
    derivative_mode: str = "reference",
    derivative_kkt_refinement_iters: int = 4,
    derivative_kkt_delta: float = 3e-6,
) -> CompiledOSQP:
    """Compile fixed-pattern batched OSQP on the HIP backend (``solver.py:685-737``).

    ``specialize`` (HIP only): shared-factor solves of a pattern with a
    generated straight-line kernel (``warpmpc_amd.codegen``) run that kernel.
    ``"auto"`` attaches the code object when one was built for this pattern,
    ``True`` builds it first if needed (hipcc), ``False`` keeps the generic
    kernel.

    ``derivative_mode`` (with ``derivatives=True``): ``"reference"`` runs the
    reference's adjoint, the eps-regularised 2*half-square adjoint matrix
    factored per instance and ``derivative_refinement_iters`` refinement
    solves; ``"kkt"`` differentiates the active-set KKT system -- the limit
    that refinement converges to wherever the rows with a positive multiplier
    are feasible -- by one factorization of a symmetric system with the
    forward KKT's pattern and ``derivative_kkt_refinement_iters`` refinement
    solves (see ``wmpc_adjoint_kkt_plan_create``).  ``derivative_kkt_delta``
    regularises the constraint diagonal of the per-instance equilibrated
    system for the float32 factorization: each refinement step contracts the
    error by about delta * |K⁻¹| of the equilibrated matrix."""
    dtype = np.dtype(dtype)
    if dtype != np.dtype(np.float32):
        raise ValueError("the fused HIP OSQP kernels compute in float32 (float64 is available for the QDLDL "
                         "factor / solve through compile_qdldl_variant)")
    settings = plan.settings
    qd = compile_qdldl_variant(
        plan.qdldl_plan,
        dtype=dtype,
        backend=qdldl_backend,
        factor_backend=qdldl_factor_backend,
        solve_backend=qdldl_solve_backend,
        transpose_work=transpose_work,
        segmented=segmented,
        segment_budget=segment_budget,
        segment_strategy=segment_strategy,
        level_scheduled_solve=level_scheduled_solve,
        level_scheduled_solve_threshold=level_scheduled_solve_threshold,
    )
    if derivative_mode not in ("reference", "kkt"):
        raise ValueError("derivative_mode must be 'reference' or 'kkt'")
    if derivatives and plan.derivative_plan is None:
        raise ValueError(
            "call build_osqp_plan(..., derivatives=True) before compile_osqp(..., derivatives=True)"
        )
    require_gpu()
    dev = DeviceOSQPPlan(plan, qd.extras["device_plan"])
    if specialize not in (True, False, "auto"):
        raise ValueError("specialize must be True, False or 'auto'")
    if specialize:
        from .. import codegen

        path = codegen.build_admm_kernel(plan) if specialize is True else codegen.find_admm_kernel(plan)
        if path is not None:
            dev.specialize(path.read_bytes())
    lib = _lib.load()
    th = torch()
    n, m = plan.n, plan.m
    nk = plan.qdldl_plan.n
    nnz_l = plan.qdldl_plan.nnz_l

    def _matrices(p_values, a_values):
        pv = _batched(p_values, plan.nnz_p, "p_values")
        av = _batched(a_values, plan.nnz_a, "a_values")
        mb = max(pv.shape[0], av.shape[0])
        if pv.shape[0] not in (1, mb) or av.shape[0] not in (1, mb):
            raise ValueError("p_values/a_values batches must be 1 or equal")
        return pv, av, mb

    def factor(p_values, a_values):
        """(lx (nnz_l, Bm), dinv (n_kkt, Bm)) in the symbolic layout (``solver.py:799-804``)."""
        pv, av, mb = _matrices(p_values, a_values)
        pm_ld = 0 if mb == 1 else mb
        p_sym = _symbolic(pv, mb)
        a_sym = _symbolic(av, mb)
        lx = th.empty((nnz_l, mb), dtype=th.float32, device=pv.device)
        dinv = th.empty((nk, mb), dtype=th.float32, device=pv.device)
        _lib.check(lib.wmpc_osqp_factor_f32(dev.handle, ptr(p_sym), ptr(a_sym), pm_ld, ptr(lx), ptr(dinv),
                                            mb, mb, stream_ptr()), "wmpc_osqp_factor_f32")
        return lx, dinv

    def _run(p_values, a_values, q, l, u, warm=None, factors=None):
        pv, av, mb = _matrices(p_values, a_values)
        q = _batched(q, n, "q")
        l = _batched(l, m, "l")
        u = _batched(u, m, "u")
        b = max(mb, q.shape[0], l.shape[0], u.shape[0])
        if warm is not None:
            warm = [_batched(w, wd, nm) for w, wd, nm in zip(warm, (n, m, m), ("warm_start_x", "warm_start_z", "warm_start_y"))]
            b = max(b, *(w.shape[0] for w in warm))
        for t_, nm in ((q, "q"), (l, "l"), (u, "u")):
            if t_.shape[0] not in (1, b):
                raise ValueError(f"{nm} batch must be 1 or {b}")
        if mb not in (1, b):
            raise ValueError(f"matrix batch must be 1 or {b}")
        dev_ = q.device
        shared = mb == 1
        pm_ld = 0 if shared else b
        p_sym = _symbolic(pv, 1 if shared else b)
        a_sym = _symbolic(av, 1 if shared else b)
        q_s, l_s, u_s = _symbolic(q, b), _symbolic(l, b), _symbolic(u, b)
        x0 = z0 = y0 = None
        if warm is not None:
            x0, z0, y0 = (_symbolic(w, b) for w in warm)
        x = th.empty((n, b), dtype=th.float32, device=dev_)
        z = th.empty((m, b), dtype=th.float32, device=dev_)
        y = th.empty((m, b), dtype=th.float32, device=dev_)
        prim = th.empty((b,), dtype=th.float32, device=dev_)
        dual = th.empty((b,), dtype=th.float32, device=dev_)
        obj = th.empty((b,), dtype=th.float32, device=dev_)
        lx = dinv = None
        f_ld = 0
        if factors is not None:
            lx = as_device_f32(factors[0], "lx")
            dinv = as_device_f32(factors[1], "dinv")
            fb = int(lx.shape[1])
            if lx.shape[0] != nnz_l or dinv.shape != (nk, fb) or fb not in (1, b):
                raise ValueError("factor shapes do not match the plan / batch")
            f_ld = 0 if (fb == 1 and b != 1) else b
            mode = 0
        else:
            mode = 1 if shared else 2
        work_len = int(lib.wmpc_osqp_work_floats(dev.handle, mode, b))
        work = th.empty((max(work_len, 1),), dtype=th.float32, device=dev_)
        _lib.check(
            lib.wmpc_osqp_solve_f32(
                dev.handle, ptr(p_sym), ptr(a_sym), pm_ld, ptr(q_s), ptr(l_s), ptr(u_s), ptr(x0), ptr(z0), ptr(y0),
                ptr(lx), ptr(dinv), f_ld, ptr(work), work.numel(), ptr(x), ptr(z), ptr(y), ptr(prim),
                ptr(dual), ptr(obj), int(settings.max_iter), b, b, stream_ptr()),
            "wmpc_osqp_solve_f32")
        return x.t().contiguous(), z.t().contiguous(), y.t().contiguous(), prim, dual, obj

    def solve(p_values, a_values, q, l, u):
        return _run(p_values, a_values, q, l, u)

    def solve_warm_start(p_values, a_values, q, l, u, warm_start_x, warm_start_z, warm_start_y):
        return _run(p_values, a_values, q, l, u, warm=(warm_start_x, warm_start_z, warm_start_y))

    def solve_with_factor(lx, dinv, p_values, a_values, q, l, u):
        return _run(p_values, a_values, q, l, u, factors=(lx, dinv))

    def solve_with_factor_warm_start(lx, dinv, p_values, a_values, q, l, u, warm_start_x, warm_start_z,
                                     warm_start_y):
        return _run(p_values, a_values, q, l, u, warm=(warm_start_x, warm_start_z, warm_start_y),
                    factors=(lx, dinv))

    def linear_solve_loop(lx, dinv, rhs):
        """``max_iter`` chained solves x <- K⁻¹ x (``solver.py:908-919``); rhs (B, n_kkt)."""
        r = _batched(rhs, nk, "rhs").t().contiguous()
        for _ in range(settings.max_iter):
            r = qd.solve(lx, dinv, r)
        return r.t().contiguous()

    def init_warm_start(batch_size: int) -> OSQPWarmStart:
        kw = dict(dtype=th.float32, device="cuda")
        return OSQPWarmStart(x=th.zeros((batch_size, n), **kw), z=th.zeros((batch_size, m), **kw),
                             y=th.zeros((batch_size, m), **kw))

    compiled = dict(
        plan=plan,
        settings=settings,
        qdldl=qd,
        solve=solve,
        factor=factor,
        solve_with_factor=solve_with_factor,
        variant=qd.variant,
        dtype=dtype,
        linear_solve_loop=linear_solve_loop,
        init_warm_start=init_warm_start,
        solve_warm_start=solve_warm_start,
        solve_with_factor_warm_start=solve_with_factor_warm_start,
        extras={"device_plan": dev},
    )
    if derivatives:
        from .adjoint import attach_derivatives

        if derivative_mode == "kkt":
            compiled = attach_derivatives(compiled, plan, derivative_kkt_refinement_iters, form="kkt",
                                          kkt_device_plan=qd.extras["device_plan"], delta=derivative_kkt_delta)
        else:
            compiled = attach_derivatives(compiled, plan, derivative_refinement_iters)
    return CompiledOSQP(**compiled)


__all__ = ["compile_osqp"]
