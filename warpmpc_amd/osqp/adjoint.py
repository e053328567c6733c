"""OSQP adjoint derivatives on the HIP backend (``compile_osqp(..., derivatives=True)``).

Mirrors the derivative half of ``compile_osqp`` (``warpmpc/jax_osqp/solver.py:
984-1474``): ``adjoint_derivative_compute`` / ``_get_vec`` / ``_get_mat`` and the
differentiable ``solve``, ``solve_warm_start``, ``solve_primal`` and ``solve_xy``.
The reference wraps them in ``jax.custom_vjp``; here they are
``torch.autograd.Function`` s whose backward is one launch of the fused adjoint
kernel (``wmpc_osqp_adjoint_f32``: assembly and LDLᵀ factorization of the
eps-regularised adjoint matrix, the solve and the iterative refinement, and the
gradients dq, dl, du, dP, dA).

Conventions are the reference's: ``dy`` is the cotangent of the split
multipliers (y_l, y_u), i.e. of |y| (rows with y < 0 take the cotangent with the
sign of -y); dP is returned on P's upper-triangle entries, off-diagonal entries
at one half of the derivative with respect to the stored value.
"""

from __future__ import annotations

import ctypes
from typing import NamedTuple

import numpy as np

from .. import _lib
from .._device import as_device_f32, ptr, stream_ptr, torch
from ..qdldl.core import DevicePlan
from .types import FixedOSQPPlan


class AdjointState(NamedTuple):
    """``adjoint_derivative_compute``'s result (``solver.py:1160``), batch-first."""

    sol: object
    y_l: object
    y_u: object
    ryl: object
    ryu: object


class DeviceAdjointPlan:
    """Owns the ``wmpc_adjoint_plan`` of a plan's derivative.

    ``form="reference"``: the reference's adjoint matrix and its own QDLDL
    plan (``wmpc_adjoint_plan_create``); ``form="kkt"``: the KKT form on the
    forward KKT pattern (``wmpc_adjoint_kkt_plan_create``), which references
    ``kkt_device_plan``, the forward KKT's QDLDL device plan."""

    def __init__(self, plan: FixedOSQPPlan, form: str = "reference", kkt_device_plan=None,
                 delta: float | None = None):
        dp = plan.derivative_plan
        if dp is None:
            raise ValueError("plan has no derivative plan (build_osqp_plan(..., derivatives=True))")
        if form not in ("reference", "kkt"):
            raise ValueError("form must be 'reference' or 'kkt'")
        lib = _lib.load()
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        self.form = form
        handle = ctypes.c_void_p()
        if form == "kkt":
            if kkt_device_plan is None:
                kkt_device_plan = DevicePlan(plan.qdldl_plan)
            self.qdldl = kkt_device_plan
            kkt = plan.kkt_upper
            if not (np.array_equal(kkt.indptr, plan.qdldl_plan.a_indptr)
                    and np.array_equal(kkt.indices, plan.qdldl_plan.a_indices)):
                raise ValueError("plan.kkt_upper is not the pattern of plan.qdldl_plan")
            self.keep = {
                "kkt_rows": i32(kkt.indices),
                "kkt_cols": i32(np.repeat(np.arange(kkt.shape[1]), np.diff(kkt.indptr))),
                "p_to_kkt": i32(plan.p_to_kkt), "p_diag_sigma": f64(plan.p_diag_sigma),
                "a_to_kkt": i32(plan.a_to_kkt), "lower_indices": i32(dp.lower_indices),
                "upper_indices": i32(dp.upper_indices), "eq_indices": i32(dp.eq_indices),
                "p_upper_rows": i32(dp.p_upper_rows), "p_upper_cols": i32(dp.p_upper_cols),
                "a_rows": i32(plan.a_rows), "a_cols": i32(plan.a_cols),
            }
            d = _lib.AdjointKktDesc()
            d.n, d.m, d.nnz_p, d.nnz_a, d.nnz_kkt = plan.n, plan.m, plan.nnz_p, plan.nnz_a, kkt.nnz
            d.n_l, d.n_u, d.n_eq = dp.n_ineq_l, dp.n_ineq_u, dp.n_eq
            d.delta = float(dp.eps if delta is None else delta)
            for name, arr in self.keep.items():
                setattr(d, name, _lib.f64p(arr) if arr.dtype == np.float64 else _lib.i32p(arr))
            self.desc = d
            _lib.check(lib.wmpc_adjoint_kkt_plan_create(ctypes.byref(d), self.qdldl.handle, ctypes.byref(handle)),
                       "wmpc_adjoint_kkt_plan_create")
            self.handle = handle
            self.dim = int(dp.adjoint_dim)
            return
        self.qdldl = DevicePlan(dp.qdldl_plan)
        adj = dp.adjoint_upper
        self.keep = {
            "adj_indptr": i32(adj.indptr), "adj_indices": i32(adj.indices), "adj_data": f64(adj.data),
            "lower_indices": i32(dp.lower_indices), "upper_indices": i32(dp.upper_indices),
            "eq_indices": i32(dp.eq_indices), "p_upper_rows": i32(dp.p_upper_rows),
            "p_upper_cols": i32(dp.p_upper_cols), "p_full_value_map": i32(dp.p_full_value_map),
            "g_rows": i32(dp.g_rows), "g_cols": i32(dp.g_cols), "g_a_value_map": i32(dp.g_a_value_map),
            "g_sign": f64(dp.g_sign), "aeq_a_value_map": i32(dp.aeq_a_value_map),
            "diag_first_pos": i32(dp.diag_first_pos), "diag_second_pos": i32(dp.diag_second_pos),
            "p_block_pos": i32(dp.p_block_pos), "g_block_pos": i32(dp.g_block_pos),
            "aeq_block_pos": i32(dp.aeq_block_pos), "glambda_block_pos": i32(dp.glambda_block_pos),
            "slack_diag_pos": i32(dp.slack_diag_pos), "aeq_t_block_pos": i32(dp.aeq_t_block_pos),
            "a_rows": i32(plan.a_rows), "a_cols": i32(plan.a_cols),
        }
        d = _lib.AdjointDesc()
        d.n, d.m = dp.n, dp.m
        d.n_l, d.n_u, d.n_eq = dp.n_ineq_l, dp.n_ineq_u, dp.n_eq
        d.nnz_adj, d.nnz_p, d.nnz_a = adj.nnz, plan.nnz_p, plan.nnz_a
        d.n_p_full = int(np.asarray(dp.p_full_value_map).size)
        d.n_g = int(np.asarray(dp.g_rows).size)
        d.n_aeq = int(np.asarray(dp.aeq_a_value_map).size)
        d.eps = float(dp.eps)
        for name, arr in self.keep.items():
            setattr(d, name, _lib.f64p(arr) if arr.dtype == np.float64 else _lib.i32p(arr))
        self.desc = d
        _lib.check(lib.wmpc_adjoint_plan_create(ctypes.byref(d), self.qdldl.handle, ctypes.byref(handle)),
                   "wmpc_adjoint_plan_create")
        self.handle = handle
        self.dim = int(dp.adjoint_dim)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().wmpc_adjoint_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


def _as_batched(x, width: int, name: str):
    x = as_device_f32(x, name)
    if x.dim() == 1:
        x = x[None, :]
    if x.dim() != 2 or x.shape[1] != width:
        raise ValueError(f"expected width {width} for {name}, got shape {tuple(x.shape)}")
    return x


def _sym(x, batch: int):
    """(1|B, w) batch-first -> contiguous (w, B) batch-major."""
    if x.shape[0] != batch:
        x = x.expand(batch, x.shape[1])
    return x.t().contiguous()


def _match_broadcast_gradient(grad, reference):
    """Sum a per-instance gradient over the batch where the input was broadcast
    (``solver.py:1233-1236``; 1-D inputs get a 1-D gradient)."""
    if reference.dim() == 1:
        return grad.sum(dim=0)
    if reference.shape[0] == 1 and grad.shape[0] != 1:
        return grad.sum(dim=0, keepdim=True)
    return grad


def attach_derivatives(compiled: dict, plan: FixedOSQPPlan, refinement_iters: int, form: str = "reference",
                       kkt_device_plan=None, delta: float | None = None) -> dict:
    """Adds the adjoint callables and the differentiable solves to ``compiled``
    (``form``: the adjoint plan kind of ``DeviceAdjointPlan``)."""
    if refinement_iters < 0:
        raise ValueError("derivative_refinement_iters must be nonnegative")
    th = torch()
    lib = _lib.load()
    dev = DeviceAdjointPlan(plan, form=form, kkt_device_plan=kkt_device_plan, delta=delta)
    n, m = plan.n, plan.m
    nnz_p, nnz_a = plan.nnz_p, plan.nnz_a
    dp_ = plan.derivative_plan
    half = dp_.half_dim
    refine = int(refinement_iters)
    p_rows = th.as_tensor(np.asarray(dp_.p_upper_rows, dtype=np.int64))
    p_cols = th.as_tensor(np.asarray(dp_.p_upper_cols, dtype=np.int64))
    a_rows = th.as_tensor(np.asarray(plan.a_rows, dtype=np.int64))
    a_cols = th.as_tensor(np.asarray(plan.a_cols, dtype=np.int64))
    pf_rows = th.as_tensor(np.asarray(plan.p_full_rows, dtype=np.int64))
    pf_cols = th.as_tensor(np.asarray(plan.p_full_cols, dtype=np.int64))
    pf_map = th.as_tensor(np.asarray(plan.p_full_value_map, dtype=np.int64))
    p_diag = np.asarray(dp_.p_upper_rows) == np.asarray(dp_.p_upper_cols)
    direct_p_scale = th.as_tensor(np.where(p_diag, 0.5, 1.0).astype(np.float32))
    solve_fwd = compiled["solve"]
    solve_warm_fwd = compiled["solve_warm_start"]

    def _idx(t, like):
        return t.to(like.device)

    def fused(p_values, a_values, l, u, x, y, dx, dy, want_sol=False, want_mat=True):
        """One launch; batch-first results (dq, dl, du, dp, da, sol)."""
        pv = _as_batched(p_values, nnz_p, "p_values")
        av = _as_batched(a_values, nnz_a, "a_values")
        l = _as_batched(l, m, "l")
        u = _as_batched(u, m, "u")
        x = _as_batched(x, n, "x")
        y = _as_batched(y, m, "y")
        dx = th.zeros_like(x) if dx is None else _as_batched(dx, n, "dx")
        dy = th.zeros_like(y) if dy is None else _as_batched(dy, m, "dy")
        ins = (pv, av, l, u, x, y, dx, dy)
        b = max(t.shape[0] for t in ins)
        for t_, nm in zip(ins, ("p_values", "a_values", "l", "u", "x", "y", "dx", "dy")):
            if t_.shape[0] not in (1, b):
                raise ValueError(f"{nm} batch must be 1 or {b}")
        shared = pv.shape[0] == 1 and av.shape[0] == 1
        pm_ld = 0 if shared else b
        p_s, a_s = _sym(pv, 1 if shared else b), _sym(av, 1 if shared else b)
        l_s, u_s, x_s, y_s, dx_s, dy_s = (_sym(t, b) for t in (l, u, x, y, dx, dy))
        dv = x.device
        kw = dict(dtype=th.float32, device=dv)
        dq = th.empty((n, b), **kw)
        dl = th.empty((m, b), **kw)
        du = th.empty((m, b), **kw)
        dpm = th.empty((nnz_p, b), **kw) if want_mat else None
        dam = th.empty((nnz_a, b), **kw) if want_mat else None
        sol = th.empty((dev.dim, b), **kw) if want_sol else None
        work_len = int(lib.wmpc_adjoint_work_floats(dev.handle, b))
        work = th.empty((max(work_len, 1),), **kw)
        _lib.check(lib.wmpc_osqp_adjoint_f32(
            dev.handle, ptr(p_s), ptr(a_s), pm_ld, ptr(l_s), ptr(u_s), ptr(x_s), ptr(y_s), ptr(dx_s), ptr(dy_s),
            refine, ptr(work), work.numel(), ptr(sol), ptr(dq), ptr(dl), ptr(du), ptr(dpm), ptr(dam), b, b,
            stream_ptr()), "wmpc_osqp_adjoint_f32")
        t = lambda z: None if z is None else z.t().contiguous()  # noqa: E731
        return t(dq), t(dl), t(du), t(dpm), t(dam), t(sol)

    def adjoint_derivative_compute(p_values, a_values, l, u, x, y, dx=None, dy=None) -> AdjointState:
        """``_adjoint_derivative_compute_wrapper`` (``solver.py:1202-1220``)."""
        _, dl, du, _, _, sol = fused(p_values, a_values, l, u, x, y, dx, dy, want_sol=True, want_mat=False)
        yb = _as_batched(y, m, "y")
        b = sol.shape[0]
        y_u = th.clamp_min(yb, 0.0).expand(b, m)
        y_l = (-th.clamp_max(yb, 0.0)).expand(b, m)
        return AdjointState(sol=sol, y_l=y_l, y_u=y_u, ryl=dl, ryu=-du)

    def adjoint_derivative_get_vec(state: AdjointState):
        """(dq, dl, du) = (rx, ryl, -ryu) (``solver.py:1162-1168``)."""
        return state.sol[:, half:half + n], state.ryl, -state.ryu

    def adjoint_derivative_get_mat(x, y, state: AdjointState):
        """(dP, dA) (``solver.py:1170-1192``)."""
        x = _as_batched(x, n, "x")
        rx = state.sol[:, half:half + n]
        b = max(x.shape[0], rx.shape[0], state.ryl.shape[0])
        x, rx = x.expand(b, n), rx.expand(b, n)
        y_l, y_u = state.y_l.expand(b, m), state.y_u.expand(b, m)
        ryl, ryu = state.ryl.expand(b, m), state.ryu.expand(b, m)
        pr, pc = _idx(p_rows, x), _idx(p_cols, x)
        ar, ac = _idx(a_rows, x), _idx(a_cols, x)
        d_p = 0.5 * (rx[:, pr] * x[:, pc] + rx[:, pc] * x[:, pr])
        d_a = (y_u[:, ar] - y_l[:, ar]) * rx[:, ac] + (ryu[:, ar] - ryl[:, ar]) * x[:, ac]
        return d_p, d_a

    def p_matvec(p_values, x):
        """P x with the full symmetric P (``solver.py:795-797``)."""
        pv = p_values.expand(x.shape[0], nnz_p)
        vals = pv[:, _idx(pf_map, x)] * x[:, _idx(pf_cols, x)]
        out = th.zeros_like(x)
        return out.index_add_(1, _idx(pf_rows, x), vals)

    def grads(p_values, a_values, q, l, u, x, y, x_bar, y_bar, obj_bar=None):
        b = x.shape[0]
        dx_total = x_bar
        if obj_bar is not None:
            pvb = _as_batched(p_values, nnz_p, "p_values")
            qb = _as_batched(q, n, "q").expand(b, n)
            dx_total = x_bar + obj_bar[:, None] * (p_matvec(pvb, x) + qb)
        dq, dl, du, d_p, d_a, _ = fused(p_values, a_values, l, u, x, y, dx_total, y_bar)
        if obj_bar is not None:
            pr, pc = _idx(p_rows, x), _idx(p_cols, x)
            d_p = d_p + obj_bar[:, None] * _idx(direct_p_scale, x)[None, :] * (x[:, pr] * x[:, pc])
            dq = dq + obj_bar[:, None] * x
        return (_match_broadcast_gradient(d_p, p_values), _match_broadcast_gradient(d_a, a_values),
                _match_broadcast_gradient(dq, q), _match_broadcast_gradient(dl, l),
                _match_broadcast_gradient(du, u))

    class _SolvePrimal(th.autograd.Function):
        @staticmethod
        def forward(ctx, p_values, a_values, q, l, u):
            x, _, y, _, _, _ = solve_fwd(p_values, a_values, q, l, u)
            ctx.save_for_backward(p_values, a_values, q, l, u, x, y)
            return x

        @staticmethod
        def backward(ctx, x_bar):
            p_values, a_values, q, l, u, x, y = ctx.saved_tensors
            return grads(p_values, a_values, q, l, u, x, y, x_bar, th.zeros_like(y))

    class _SolveXY(th.autograd.Function):
        @staticmethod
        def forward(ctx, p_values, a_values, q, l, u):
            x, _, y, _, _, _ = solve_fwd(p_values, a_values, q, l, u)
            ctx.save_for_backward(p_values, a_values, q, l, u, x, y)
            return x, y

        @staticmethod
        def backward(ctx, x_bar, y_bar):
            p_values, a_values, q, l, u, x, y = ctx.saved_tensors
            return grads(p_values, a_values, q, l, u, x, y, x_bar, y_bar)

    class _Solve(th.autograd.Function):
        @staticmethod
        def forward(ctx, p_values, a_values, q, l, u):
            out = solve_fwd(p_values, a_values, q, l, u)
            x, _, y, prim, dual, _ = out
            ctx.mark_non_differentiable(prim, dual)
            ctx.save_for_backward(p_values, a_values, q, l, u, x, y)
            return out

        @staticmethod
        def backward(ctx, x_bar, z_bar, y_bar, prim_bar, dual_bar, obj_bar):
            p_values, a_values, q, l, u, x, y = ctx.saved_tensors
            return grads(p_values, a_values, q, l, u, x, y, x_bar, y_bar, obj_bar)

    class _SolveWarm(th.autograd.Function):
        @staticmethod
        def forward(ctx, p_values, a_values, q, l, u, wx, wz, wy):
            out = solve_warm_fwd(p_values, a_values, q, l, u, wx, wz, wy)
            x, _, y, prim, dual, _ = out
            ctx.mark_non_differentiable(prim, dual)
            ctx.save_for_backward(p_values, a_values, q, l, u, x, y)
            return out

        @staticmethod
        def backward(ctx, x_bar, z_bar, y_bar, prim_bar, dual_bar, obj_bar):
            p_values, a_values, q, l, u, x, y = ctx.saved_tensors
            # the warm start gets zero cotangents (solver.py:1464-1466)
            return (*grads(p_values, a_values, q, l, u, x, y, x_bar, y_bar, obj_bar), None, None, None)

    def _inputs(p_values, a_values, q, l, u):
        return (as_device_f32(p_values, "p_values"), as_device_f32(a_values, "a_values"),
                as_device_f32(q, "q"), as_device_f32(l, "l"), as_device_f32(u, "u"))

    def solve_primal(p_values, a_values, q, l, u):
        return _SolvePrimal.apply(*_inputs(p_values, a_values, q, l, u))

    def solve_xy(p_values, a_values, q, l, u):
        return _SolveXY.apply(*_inputs(p_values, a_values, q, l, u))

    def solve(p_values, a_values, q, l, u):
        return _Solve.apply(*_inputs(p_values, a_values, q, l, u))

    def solve_warm_start(p_values, a_values, q, l, u, warm_start_x, warm_start_z, warm_start_y):
        return _SolveWarm.apply(*_inputs(p_values, a_values, q, l, u), warm_start_x, warm_start_z,
                                warm_start_y)

    out = dict(compiled)
    out.update(
        solve=solve,
        solve_warm_start=solve_warm_start,
        solve_primal=solve_primal,
        solve_xy=solve_xy,
        adjoint_derivative_compute=adjoint_derivative_compute,
        adjoint_derivative_get_vec=adjoint_derivative_get_vec,
        adjoint_derivative_get_mat=adjoint_derivative_get_mat,
    )
    extras = dict(out.get("extras", {}))
    extras.update(adjoint_plan=dev, adjoint_fused=fused)
    out["extras"] = extras
    return out


__all__ = ["AdjointState", "DeviceAdjointPlan", "attach_derivatives"]
