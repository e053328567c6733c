"""ctypes front end of the CPU oracle (``oracle/liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` as the checker.  Product code never imports
this module.

Array conventions are instance-major ``[batch, width]`` (C-contiguous), i.e. the
reference's "batch" layout; callers transpose symbolic-layout data themselves.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB = None

_i32p = ctypes.POINTER(ctypes.c_int32)


class _Pattern(ctypes.Structure):
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


_f64p = ctypes.POINTER(ctypes.c_double)


class _OsqpDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("m", ctypes.c_int32),
        ("nnz_p", ctypes.c_int32),
        ("nnz_a", ctypes.c_int32),
        ("nnz_kkt", ctypes.c_int32),
        ("max_iter", ctypes.c_int32),
        ("nnz_p_full", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("sigma", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("kkt", ctypes.POINTER(_Pattern)),
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


class _AdjDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32), ("m", ctypes.c_int32), ("n_l", ctypes.c_int32), ("n_u", ctypes.c_int32),
        ("n_eq", ctypes.c_int32), ("dim", ctypes.c_int32), ("nnz_adj", ctypes.c_int32),
        ("nnz_p", ctypes.c_int32), ("nnz_a", ctypes.c_int32), ("refine", ctypes.c_int32),
        ("eps", ctypes.c_double),
        ("kkt", ctypes.POINTER(_Pattern)),
        ("adj_rows", _i32p), ("adj_cols", _i32p), ("adj_base", _f64p),
        ("lower", _i32p), ("upper", _i32p), ("eq", _i32p),
        ("n_pf", ctypes.c_int32), ("n_g", ctypes.c_int32), ("n_e", ctypes.c_int32),
        ("p_full_map", _i32p), ("g_rows", _i32p), ("g_cols", _i32p), ("g_map", _i32p), ("g_sign", _f64p),
        ("e_map", _i32p), ("diag_first_pos", _i32p), ("diag_second_pos", _i32p), ("p_pos", _i32p),
        ("g_pos", _i32p), ("aeq_pos", _i32p), ("glambda_pos", _i32p), ("slack_pos", _i32p),
        ("aeq_t_pos", _i32p), ("p_rows", _i32p), ("p_cols", _i32p), ("a_rows", _i32p), ("a_cols", _i32p),
    ]


def build() -> pathlib.Path:
    """Compile ``liboracle.so`` with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _HERE / "liboracle.so"


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        _LIB = ctypes.CDLL(str(path))
        vp = ctypes.c_void_p
        i64 = ctypes.c_int64
        for sfx in ("f64", "f32"):
            getattr(_LIB, f"oracle_qdldl_factor_{sfx}").argtypes = [ctypes.POINTER(_Pattern), i64, vp, vp, vp, vp]
            getattr(_LIB, f"oracle_qdldl_solve_{sfx}").argtypes = [
                ctypes.POINTER(_Pattern), i64, vp, i64, vp, i64, vp, vp]
            getattr(_LIB, f"oracle_osqp_solve_{sfx}").argtypes = [
                ctypes.POINTER(_OsqpDesc), i64, vp, i64, vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
            getattr(_LIB, f"oracle_adjoint_{sfx}").argtypes = [
                ctypes.POINTER(_AdjDesc), i64, vp, i64, vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                vp, vp]
        _LIB.oracle_num_threads.restype = ctypes.c_int
        _LIB.oracle_set_num_threads.argtypes = [ctypes.c_int]
        _LIB.oracle_set_num_threads.restype = None
    return _LIB


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class OraclePattern:
    """Keeps the int32 tables alive next to the ctypes struct that points at them."""

    def __init__(self, plan):
        from warpmpc_amd.qdldl.tables import ragged_program

        self.plan = plan
        self.tables = ragged_program(plan)
        s = _Pattern()
        s.n = plan.n
        s.nnz_a = plan.nnz_a
        s.nnz_l = plan.nnz_l
        for name, arr in self.tables.items():
            setattr(s, name, arr.ctypes.data_as(_i32p))
        self.struct = s


def _dtype_sfx(dtype) -> str:
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        return "f64"
    if dtype == np.float32:
        return "f32"
    raise ValueError(f"oracle supports float32/float64, got {dtype}")


def qdldl_factor(plan, a_values: np.ndarray, dtype=np.float64):
    """Factor ``a_values[batch, nnz_a]`` -> ``(lx[batch, nnz_l], d[batch, n], dinv[batch, n])``."""
    sfx = _dtype_sfx(dtype)
    pat = OraclePattern(plan)
    a = np.ascontiguousarray(np.atleast_2d(a_values), dtype=dtype)
    b = a.shape[0]
    lx = np.zeros((b, plan.nnz_l), dtype=dtype)
    d = np.zeros((b, plan.n), dtype=dtype)
    dinv = np.zeros((b, plan.n), dtype=dtype)
    rc = getattr(lib(), f"oracle_qdldl_factor_{sfx}")(ctypes.byref(pat.struct), b, _ptr(a), _ptr(lx), _ptr(d), _ptr(dinv))
    if rc:
        raise RuntimeError("oracle factor failed")
    return lx, d, dinv


def qdldl_solve(plan, lx: np.ndarray, dinv: np.ndarray, rhs: np.ndarray, dtype=np.float64):
    """Solve with factors ``lx[b or 1, nnz_l]``, ``dinv[b or 1, n]`` for ``rhs[batch, n]``."""
    sfx = _dtype_sfx(dtype)
    pat = OraclePattern(plan)
    lx = np.ascontiguousarray(np.atleast_2d(lx), dtype=dtype)
    dinv = np.ascontiguousarray(np.atleast_2d(dinv), dtype=dtype)
    rhs = np.ascontiguousarray(np.atleast_2d(rhs), dtype=dtype)
    b = rhs.shape[0]
    out = np.zeros_like(rhs)
    lxs = 0 if lx.shape[0] == 1 else plan.nnz_l
    ds = 0 if dinv.shape[0] == 1 else plan.n
    rc = getattr(lib(), f"oracle_qdldl_solve_{sfx}")(
        ctypes.byref(pat.struct), b, _ptr(lx), lxs, _ptr(dinv), ds, _ptr(rhs), _ptr(out))
    if rc:
        raise RuntimeError("oracle solve failed")
    return out


class OracleOSQP:
    """Fixed-iteration OSQP oracle bound to one :class:`FixedOSQPPlan`."""

    def __init__(self, plan, max_iter: int | None = None):
        self.plan = plan
        self.pattern = OraclePattern(plan.qdldl_plan)
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
        d = _OsqpDesc()
        d.n, d.m = plan.n, plan.m
        d.nnz_p, d.nnz_a, d.nnz_kkt = plan.nnz_p, plan.nnz_a, plan.nnz_kkt
        d.max_iter = plan.settings.max_iter if max_iter is None else int(max_iter)
        d.nnz_p_full = int(plan.p_full_rows.size)
        d.sigma = plan.settings.sigma
        d.alpha = plan.settings.alpha
        d.kkt = ctypes.pointer(self.pattern.struct)
        for name, arr in self.keep.items():
            ptype = _f64p if arr.dtype == np.float64 else _i32p
            setattr(d, name, arr.ctypes.data_as(ptype))
        self.desc = d

    def solve(self, p_values, a_values, q, l, u, x0=None, z0=None, y0=None, dtype=np.float64):
        """Returns ``(x, z, y, prim_res, dual_res, obj_val)`` like ``CompiledOSQP.solve``."""
        sfx = _dtype_sfx(dtype)
        plan = self.plan
        pv = np.ascontiguousarray(np.atleast_2d(p_values), dtype=dtype)
        av = np.ascontiguousarray(np.atleast_2d(a_values), dtype=dtype)
        q = np.ascontiguousarray(np.atleast_2d(q), dtype=dtype)
        l = np.ascontiguousarray(np.atleast_2d(l), dtype=dtype)
        u = np.ascontiguousarray(np.atleast_2d(u), dtype=dtype)
        b = max(pv.shape[0], av.shape[0], q.shape[0], l.shape[0], u.shape[0])
        q = np.ascontiguousarray(np.broadcast_to(q, (b, plan.n)))
        l = np.ascontiguousarray(np.broadcast_to(l, (b, plan.m)))
        u = np.ascontiguousarray(np.broadcast_to(u, (b, plan.m)))
        shared = pv.shape[0] == 1 and av.shape[0] == 1
        if not shared:
            pv = np.ascontiguousarray(np.broadcast_to(pv, (b, plan.nnz_p)))
            av = np.ascontiguousarray(np.broadcast_to(av, (b, plan.nnz_a)))
        ps = 0 if shared else plan.nnz_p
        as_ = 0 if shared else plan.nnz_a

        def opt(v, w):
            if v is None:
                return None, None
            arr = np.ascontiguousarray(np.broadcast_to(np.atleast_2d(np.asarray(v, dtype=dtype)), (b, w)))
            return arr, _ptr(arr)

        x0a, x0p = opt(x0, plan.n)
        z0a, z0p = opt(z0, plan.m)
        y0a, y0p = opt(y0, plan.m)
        x = np.zeros((b, plan.n), dtype=dtype)
        z = np.zeros((b, plan.m), dtype=dtype)
        y = np.zeros((b, plan.m), dtype=dtype)
        prim = np.zeros(b, dtype=dtype)
        dual = np.zeros(b, dtype=dtype)
        obj = np.zeros(b, dtype=dtype)
        rc = getattr(lib(), f"oracle_osqp_solve_{sfx}")(
            ctypes.byref(self.desc), b, _ptr(pv), ps, _ptr(av), as_, _ptr(q), _ptr(l), _ptr(u),
            x0p, z0p, y0p, _ptr(x), _ptr(z), _ptr(y), _ptr(prim), _ptr(dual), _ptr(obj))
        if rc:
            raise RuntimeError("oracle osqp failed")
        del x0a, z0a, y0a
        return x, z, y, prim, dual, obj


class OracleAdjoint:
    """OSQP adjoint derivative oracle bound to a plan built with ``derivatives=True``.

    ``compute`` returns ``(sol, ryl, ryu, dq, dl, du, dp, da)`` per instance,
    restating ``solver.py:1035-1192``; ``refine`` is
    ``derivative_refinement_iters``.
    """

    def __init__(self, plan, refine: int = 100):
        dp = plan.derivative_plan
        if dp is None:
            raise ValueError("plan has no derivative plan (build_osqp_plan(..., derivatives=True))")
        self.plan = plan
        self.dp = dp
        self.pattern = OraclePattern(dp.qdldl_plan)
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        adj = dp.adjoint_upper
        rows = np.asarray(adj.indices, dtype=np.int32)
        cols = np.repeat(np.arange(dp.adjoint_dim, dtype=np.int32), np.diff(adj.indptr))
        self.keep = {
            "adj_rows": i32(rows), "adj_cols": i32(cols), "adj_base": f64(adj.data),
            "lower": i32(dp.lower_indices), "upper": i32(dp.upper_indices), "eq": i32(dp.eq_indices),
            "p_full_map": i32(dp.p_full_value_map), "g_rows": i32(dp.g_rows), "g_cols": i32(dp.g_cols),
            "g_map": i32(dp.g_a_value_map), "g_sign": f64(dp.g_sign), "e_map": i32(dp.aeq_a_value_map),
            "diag_first_pos": i32(dp.diag_first_pos), "diag_second_pos": i32(dp.diag_second_pos),
            "p_pos": i32(dp.p_block_pos), "g_pos": i32(dp.g_block_pos), "aeq_pos": i32(dp.aeq_block_pos),
            "glambda_pos": i32(dp.glambda_block_pos), "slack_pos": i32(dp.slack_diag_pos),
            "aeq_t_pos": i32(dp.aeq_t_block_pos), "p_rows": i32(dp.p_upper_rows), "p_cols": i32(dp.p_upper_cols),
            "a_rows": i32(plan.a_rows), "a_cols": i32(plan.a_cols),
        }
        d = _AdjDesc()
        d.n, d.m = dp.n, dp.m
        d.n_l, d.n_u, d.n_eq = dp.n_ineq_l, dp.n_ineq_u, dp.n_eq
        d.dim, d.nnz_adj = dp.adjoint_dim, adj.nnz
        d.nnz_p, d.nnz_a = plan.nnz_p, plan.nnz_a
        d.refine = int(refine)
        d.eps = dp.eps
        d.kkt = ctypes.pointer(self.pattern.struct)
        d.n_pf, d.n_g, d.n_e = dp.p_full_value_map.size, dp.g_rows.size, dp.aeq_rows.size
        for name, arr in self.keep.items():
            setattr(d, name, arr.ctypes.data_as(_f64p if arr.dtype == np.float64 else _i32p))
        self.desc = d

    def compute(self, p_values, a_values, l, u, x, y, dx, dy, dtype=np.float64):
        sfx = _dtype_sfx(dtype)
        plan = self.plan
        n, m = plan.n, plan.m
        arr = lambda v, w: np.ascontiguousarray(np.atleast_2d(np.asarray(v, dtype=dtype)))  # noqa: E731
        pv, av = arr(p_values, plan.nnz_p), arr(a_values, plan.nnz_a)
        ins = [arr(v, 0) for v in (l, u, x, y, dx, dy)]
        b = max(pv.shape[0], av.shape[0], *(v.shape[0] for v in ins))
        widths = (m, m, n, m, n, m)
        ins = [np.ascontiguousarray(np.broadcast_to(v, (b, w))) for v, w in zip(ins, widths)]
        ps = 0 if pv.shape[0] == 1 else plan.nnz_p
        as_ = 0 if av.shape[0] == 1 else plan.nnz_a
        if ps:
            pv = np.ascontiguousarray(np.broadcast_to(pv, (b, plan.nnz_p)))
        if as_:
            av = np.ascontiguousarray(np.broadcast_to(av, (b, plan.nnz_a)))
        out = [np.zeros((b, w), dtype=dtype) for w in (self.dp.adjoint_dim, m, m, n, m, m, plan.nnz_p, plan.nnz_a)]
        rc = getattr(lib(), f"oracle_adjoint_{sfx}")(
            ctypes.byref(self.desc), b, _ptr(pv), ps, _ptr(av), as_, *(_ptr(v) for v in ins), *(_ptr(o) for o in out))
        if rc:
            raise RuntimeError("oracle adjoint failed")
        return tuple(out)


def adjoint_exact(plan, p_values, a_values, l, u, x, y, dx, dy, clamp_slacks=False):
    """The reference adjoint (``solver.py:1047-1192``) with its refinement
    carried to convergence, in float64 numpy.

    ``clamp_slacks``: the slack s of an inequality row with a positive
    multiplier is taken as min(s, 0) (no change where those rows are feasible,
    as at a converged solution), which makes U the Jacobian of the active-set
    KKT system at an iterate that violates such a row.

    The reference solves M0 sol = rhs, M0 = [[I, U], [Uᵀ, 0]] with
    U = [[P, Gᵀ diag(lambda), Aeqᵀ], [G, diag(s), 0], [Aeq, 0, 0]] (G, h, lambda,
    s as in the reference), iteratively refining a solve with the
    eps-perturbed matrix.  The limit of that refinement is sol = [0; U⁻¹ r]
    (a + U b = r, Uᵀ a = 0), which this function forms directly: U is built
    dense per instance and solved with ``numpy.linalg.solve``.  The outputs
    follow the reference's mapping (``solver.py:1141-1192``).  Returns
    ``(sol, ryl, ryu, dq, dl, du, dp, da)`` per instance like
    ``OracleAdjoint.compute``.
    """
    dp_ = plan.derivative_plan
    if dp_ is None:
        raise ValueError("plan has no derivative plan (build_osqp_plan(..., derivatives=True))")
    n, m = plan.n, plan.m
    f = lambda v, w: np.atleast_2d(np.asarray(v, dtype=np.float64))  # noqa: E731
    pv, av = f(p_values, plan.nnz_p), f(a_values, plan.nnz_a)
    ins = [f(v, 0) for v in (l, u, x, y, dx, dy)]
    b = max(pv.shape[0], av.shape[0], *(v.shape[0] for v in ins))
    pv = np.broadcast_to(pv, (b, plan.nnz_p))
    av = np.broadcast_to(av, (b, plan.nnz_a))
    l, u, x, y, dx, dy = (np.broadcast_to(v, (b, w)) for v, w in zip(ins, (m, m, n, m, n, m)))
    lo, up, eq = (np.asarray(i, dtype=np.int64) for i in (dp_.lower_indices, dp_.upper_indices, dp_.eq_indices))
    nl, nu, ne = lo.size, up.size, eq.size
    ni = nl + nu
    half = n + ni + ne
    pr, pc = (np.asarray(i, dtype=np.int64) for i in (dp_.p_upper_rows, dp_.p_upper_cols))
    ar, ac = np.asarray(plan.a_rows, dtype=np.int64), np.asarray(plan.a_cols, dtype=np.int64)
    outs = [np.zeros((b, w)) for w in (2 * half, m, m, n, m, m, plan.nnz_p, plan.nnz_a)]
    for k in range(b):
        pm = np.zeros((n, n))
        pm[pr, pc] = pv[k]
        pm[pc, pr] = pv[k]
        am = np.zeros((m, n))
        am[ar, ac] = av[k]
        g = np.vstack([-am[lo], am[up]])
        h = np.concatenate([-l[k, lo], u[k, up]])
        y_u = np.maximum(y[k], 0.0)
        y_l = -np.minimum(y[k], 0.0)
        lam = np.concatenate([y_l[lo], y_u[up]])
        s = g @ x[k] - h
        if clamp_slacks:
            s = np.where(lam > 0.0, np.minimum(s, 0.0), s)
        um = np.zeros((half, half))
        um[:n, :n] = pm
        um[:n, n:n + ni] = (lam[:, None] * g).T
        um[:n, n + ni:] = am[eq].T
        um[n:n + ni, :n] = g
        um[n:n + ni, n:n + ni] = np.diag(s)
        um[n + ni:, :n] = am[eq]
        yeq = y[k, eq]
        d_nu = np.where(yeq >= 0.0, dy[k, eq], -dy[k, eq])
        r = np.concatenate([-dx[k], -dy[k, lo], -dy[k, up], -d_nu])
        # an inequality row without a positive multiplier has a zero column
        # in the x rows of U: its b_lambda follows from b_x afterwards, and U
        # is solved on the other rows (U itself is singular when such a row
        # also has s = 0)
        keep = np.concatenate([np.ones(n, bool), lam > 0.0, np.ones(ne, bool)])
        bsol = np.zeros(half)
        bsol[keep] = np.linalg.solve(um[np.ix_(keep, keep)], r[keep])
        drop = np.flatnonzero(~keep)
        with np.errstate(divide="ignore", invalid="ignore"):
            bsol[drop] = (r[drop] - um[drop, :n] @ bsol[:n]) / s[drop - n]
        sol = np.concatenate([np.zeros(half), bsol])
        rx = bsol[:n]
        r_yl = np.zeros(m)
        r_yu = np.zeros(m)
        r_yl[lo] = -bsol[n:n + nl]
        r_yu[up] = bsol[n + nl:n + ni]
        es = bsol[n + ni:]
        with np.errstate(divide="ignore", invalid="ignore"):
            r_yu[eq] = np.where(yeq >= 0.0, es / yeq, 0.0)
            r_yl[eq] = np.where(yeq >= 0.0, 0.0, -es / yeq)
        ryl = -(r_yl * y_l)
        ryu = r_yu * y_u
        dpv = 0.5 * (rx[pr] * x[k, pc] + rx[pc] * x[k, pr])
        dav = (y_u[ar] - y_l[ar]) * rx[ac] + (ryu[ar] - ryl[ar]) * x[k, ac]
        for o, v in zip(outs, (sol, ryl, ryu, rx, ryl, -ryu, dpv, dav)):
            o[k] = v
    return tuple(outs)


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def set_threads(n: int) -> None:
    """Threads of the oracle's OpenMP loops from now on."""
    lib().oracle_set_num_threads(int(n))


__all__ = ["OracleAdjoint", "adjoint_exact", "OracleOSQP", "OraclePattern", "build", "lib", "num_threads", "qdldl_factor",
           "qdldl_solve", "set_threads"]
