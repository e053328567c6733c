"""Heterogeneous batches: several fixed-pattern OSQP problem sets in one launch.

The reference compiles one solver per sparsity pattern and solves each
pattern's batch with its own calls.  ``HeterogeneousOSQP`` takes compiled HIP
solvers of up to 12 patterns and solves all their batches in one persistent
launch (``wmpc_osqp_solve_hetero_f32``): wavefronts pull work items (one
wavefront of instances of one pattern) from a queue ordered largest pattern
first, each instance computing exactly what ``CompiledOSQP.solve`` /
``solve_warm_start`` compute for it (factor per instance).

``shift_warm_start`` forms the next warm start of an MPC from the previous
solution advanced one stage (``mpc_shift_maps``), as a device row gather.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import _lib
from .._device import as_device_f32, ptr, stream_ptr, torch
from .types import OSQPWarmStart


def _batched(x, width: int, name: str):
    x = as_device_f32(x, name)
    if x.dim() == 1:
        x = x[None, :]
    if x.dim() != 2 or x.shape[1] != width:
        raise ValueError(f"{name} must have shape (batch, {width}), got {tuple(x.shape)}")
    return x


def _sym(x, batch: int):
    if x.shape[0] != batch:
        x = x.expand(batch, x.shape[1])
    return x.t().contiguous()


class HeterogeneousOSQP:
    """One launch over the batches of several compiled OSQP solvers."""

    def __init__(self, solvers):
        solvers = list(solvers)
        if not 1 <= len(solvers) <= 12:
            raise ValueError("need 1 to 12 compiled solvers")
        self.solvers = solvers
        self.plans = [s.plan for s in solvers]
        self._dev = [s.extras["device_plan"] for s in solvers]  # keep the plans alive
        lib = _lib.load()
        arr = (ctypes.c_void_p * len(solvers))(*[d.handle.value for d in self._dev])
        handle = ctypes.c_void_p()
        _lib.check(lib.wmpc_osqp_hetero_plan_create(arr, len(solvers), ctypes.byref(handle)),
                   "wmpc_osqp_hetero_plan_create")
        self.handle = handle

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.load().wmpc_osqp_hetero_plan_destroy(h)
            except Exception:
                pass
            self.handle = None

    def work_floats(self, lds) -> int:
        arr = (ctypes.c_int64 * len(self.plans))(*[int(v) for v in lds])
        return int(_lib.load().wmpc_osqp_hetero_work_floats(self.handle, arr))

    def solve_symbolic(self, problems, warm_starts=None, max_iter: int | None = None, work=None):
        """Batch-major form: problems[k] = dict(p, a, q, l, u) with p/a of shape
        (nnz, B or 1) and q/l/u (width, B), contiguous float32 on the GPU;
        warm_starts[k] = (x0, z0, y0) batch-major or None.  Returns per pattern
        (x, z, y, prim, dual, obj), batch-major."""
        th = torch()
        lib = _lib.load()
        if len(problems) != len(self.plans):
            raise ValueError("one problem set per solver")
        ios = (_lib.HeteroIO * len(self.plans))()
        outs = []
        keep = []
        for k, (plan, pr) in enumerate(zip(self.plans, problems)):
            b = int(pr["q"].shape[1])
            pm_ld = 0 if pr["p"].shape[1] == 1 and b != 1 else b
            kw = dict(dtype=th.float32, device=pr["q"].device)
            o = (th.empty((plan.n, b), **kw), th.empty((plan.m, b), **kw), th.empty((plan.m, b), **kw),
                 th.empty((b,), **kw), th.empty((b,), **kw), th.empty((b,), **kw))
            outs.append(o)
            ws = None if warm_starts is None else warm_starts[k]
            io = ios[k]
            io.p_values, io.a_values, io.pm_ld = ptr(pr["p"]), ptr(pr["a"]), pm_ld
            io.q, io.l, io.u = ptr(pr["q"]), ptr(pr["l"]), ptr(pr["u"])
            if ws is not None:
                io.x0, io.z0, io.y0 = (ptr(w) for w in ws)
                keep.append(ws)
            io.x, io.z, io.y, io.prim, io.dual, io.obj = (ptr(t) for t in o)
            io.batch = io.ld = b
        need = self.work_floats([io.ld for io in ios])
        if work is None or work.numel() < need:
            work = th.empty((max(need, 1),), dtype=th.float32, device="cuda")
        iters = self.plans[0].settings.max_iter if max_iter is None else int(max_iter)
        _lib.check(lib.wmpc_osqp_solve_hetero_f32(self.handle, ios, iters, ptr(work), work.numel(), stream_ptr()),
                   "wmpc_osqp_solve_hetero_f32")
        return outs

    def solve(self, problems, warm_starts=None, max_iter: int | None = None):
        """Batch-first form, like ``CompiledOSQP.solve``: problems[k] =
        (p_values, a_values, q, l, u); warm_starts[k] = OSQPWarmStart / (x, z, y)
        or None.  Returns per pattern (x, z, y, prim_res, dual_res, obj_val)."""
        sym = []
        wsym = None if warm_starts is None else []
        for k, (plan, pr) in enumerate(zip(self.plans, problems)):
            pv = _batched(pr[0], plan.nnz_p, "p_values")
            av = _batched(pr[1], plan.nnz_a, "a_values")
            q, l, u = (_batched(v, w, nm) for v, w, nm in zip(pr[2:], (plan.n, plan.m, plan.m), "qlu"))
            b = max(pv.shape[0], av.shape[0], q.shape[0], l.shape[0], u.shape[0])
            mb = max(pv.shape[0], av.shape[0])
            shared = mb == 1 and b != 1
            sym.append(dict(p=_sym(pv, 1 if shared else b), a=_sym(av, 1 if shared else b), q=_sym(q, b),
                            l=_sym(l, b), u=_sym(u, b)))
            if warm_starts is not None:
                ws = warm_starts[k]
                if ws is None:
                    wsym.append(None)
                else:
                    wx, wz, wy = (ws.x, ws.z, ws.y) if isinstance(ws, OSQPWarmStart) else ws
                    wsym.append(tuple(_sym(_batched(v, w, nm), b) for v, w, nm in
                                      zip((wx, wz, wy), (plan.n, plan.m, plan.m), ("x0", "z0", "y0"))))
        outs = self.solve_symbolic(sym, wsym, max_iter=max_iter)
        return [tuple(t.t().contiguous() if t.dim() == 2 else t for t in o) for o in outs]


class ShiftMaps:
    """Device copies of an MPC pattern's shift maps (``mpc_shift_maps``)."""

    def __init__(self, var_map, con_map, n: int, m: int):
        th = torch()
        self.var = th.as_tensor(np.asarray(var_map, dtype=np.int32), device="cuda")
        self.con = th.as_tensor(np.asarray(con_map, dtype=np.int32), device="cuda")
        self.n, self.m = n, m


def shift_warm_start(maps: ShiftMaps, x, z, y, out=None):
    """Shifted warm start from a previous solution, batch-major (width, B)
    device arrays; returns (x0, z0, y0)."""
    th = torch()
    lib = _lib.load()
    b = int(x.shape[1])
    if out is None:
        out = (th.empty_like(x), th.empty_like(z), th.empty_like(y))
    for src, dst, rows, width in ((x, out[0], maps.var, maps.n), (z, out[1], maps.con, maps.m),
                                  (y, out[2], maps.con, maps.m)):
        if src.shape != (width, b) or dst.shape != (width, b):
            raise ValueError("shift_warm_start expects batch-major (width, B) arrays")
        _lib.check(lib.wmpc_gather_rows_f32(ptr(rows), int(rows.numel()), width, ptr(src), b, ptr(dst), b, b,
                                            stream_ptr()), "wmpc_gather_rows_f32")
    return out


__all__ = ["HeterogeneousOSQP", "ShiftMaps", "shift_warm_start"]
