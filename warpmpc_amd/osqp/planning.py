"""Host-side setup of the fixed-pattern direct-solver OSQP workspace.

Restates ``warpmpc/jax_osqp/solver.py:64-343, 574-662``:

* ``upper_with_diagonal`` — upper triangle of P with every diagonal present
  (``solver.py:64-78``);
* ``build_kkt_and_maps`` — upper CSC of ``[P + sigma I, Aᵀ; A, -diag(1/rho)]`` and
  the value maps ``p_to_kkt``, ``a_to_kkt``, ``rho_to_kkt`` (``solver.py:81-150``);
* ``bounds_type`` / ``rho_vectors`` — constraint classification and the vector
  rho (equality rows ``1e3 rho``, free rows ``1e-6``) (``solver.py:153-183``);
* ``compute_scaling`` — modified Ruiz equilibration (``solver.py:262-343``);
* ``build_osqp_plan`` (``solver.py:574-662``).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from ..qdldl.planning import build_qdldl_plan
from .types import FixedOSQPPlan, OSQPSettings

OSQP_INFTY = 1e30
OSQP_MIN_SCALING = 1e-4
OSQP_MAX_SCALING = 1e4


def upper_with_diagonal(matrix) -> sp.csc_matrix:
    """Upper triangle of P in canonical CSC (``solver.py:64-78``).

    The reference also tries to insert structurally missing diagonal entries by
    adding a sparse matrix of explicit zeros; scipy's sparse addition prunes
    zeros, so that insertion never changes the pattern.  The observable result is
    ``triu(P)`` in canonical form, which is what this returns; missing diagonals
    only appear in the KKT matrix (as ``sigma``, see ``build_kkt_and_maps``).
    """
    upper = sp.triu(sp.csc_matrix(matrix), format="csc")
    upper.sum_duplicates()
    upper.sort_indices()
    return upper


def csc_rows_cols(matrix: sp.csc_matrix) -> tuple[np.ndarray, np.ndarray]:
    cols = np.repeat(np.arange(matrix.shape[1], dtype=np.int32), np.diff(matrix.indptr))
    return np.asarray(matrix.indices, dtype=np.int32), cols


def build_kkt_and_maps(p_upper: sp.csc_matrix, a_matrix: sp.csc_matrix, settings: OSQPSettings,
                       rho_inv_vec: np.ndarray):
    """KKT upper pattern/values and the value-scatter maps (``solver.py:81-150``)."""
    n = p_upper.shape[0]
    m = a_matrix.shape[0]
    p_rows, p_cols = csc_rows_cols(p_upper)
    a_rows, a_cols = csc_rows_cols(a_matrix)
    p_vals = np.asarray(p_upper.data, dtype=np.float64).copy()
    p_is_diag = p_rows == p_cols
    p_vals[p_is_diag] += settings.sigma
    # columns of P without a stored diagonal still get sigma on the KKT diagonal
    has_diag = np.zeros(n, dtype=bool)
    has_diag[p_cols[p_is_diag]] = True
    missing = np.flatnonzero(~has_diag)
    rows = np.concatenate([p_rows, a_cols, n + np.arange(m), missing]).astype(np.int64)
    cols = np.concatenate([p_cols, n + a_rows, n + np.arange(m), missing]).astype(np.int64)
    vals = np.concatenate([p_vals, np.asarray(a_matrix.data, dtype=np.float64),
                           -np.asarray(rho_inv_vec, dtype=np.float64),
                           np.full(missing.size, settings.sigma)])
    order = np.lexsort((rows, cols))
    pos = np.empty(order.size, dtype=np.int64)
    pos[order] = np.arange(order.size)
    counts = np.bincount(cols, minlength=n + m)
    indptr = np.zeros(n + m + 1, dtype=np.int32)
    np.cumsum(counts, out=indptr[1:])
    kkt = sp.csc_matrix((vals[order], rows[order].astype(np.int32), indptr), shape=(n + m, n + m))
    nnz_p = p_rows.size
    nnz_a = a_rows.size
    p_to_kkt = pos[:nnz_p].astype(np.int32)
    a_to_kkt = pos[nnz_p : nnz_p + nnz_a].astype(np.int32)
    rho_to_kkt = pos[nnz_p + nnz_a : nnz_p + nnz_a + m].astype(np.int32)
    p_diag_sigma = np.where(p_is_diag, settings.sigma, 0.0)
    return kkt, p_to_kkt, p_diag_sigma, a_to_kkt, rho_to_kkt


def bounds_type(l: np.ndarray, u: np.ndarray) -> np.ndarray:
    """-1 free, 1 equality (|u - l| < 1e-4), 0 inequality (``solver.py:153-161``)."""
    l = np.asarray(l, dtype=float)
    u = np.asarray(u, dtype=float)
    out = np.zeros(l.shape, dtype=np.int32)
    out[(l <= -OSQP_INFTY) & (u >= OSQP_INFTY)] = -1
    out[np.abs(u - l) < 1e-4] = 1
    return out


def rho_vectors(settings: OSQPSettings, m: int, l, u):
    """Vector rho, its inverse and constraint types (``solver.py:164-183``)."""
    if settings.rho_is_vec:
        if l is None or u is None:
            raise ValueError("l and u are required when settings.rho_is_vec=True")
        constr_type = bounds_type(np.asarray(l)[:m], np.asarray(u)[:m])
        rho_vec = np.where(constr_type < 0, 1e-6,
                           np.where(constr_type > 0, 1e3 * settings.rho, settings.rho))
    else:
        constr_type = np.zeros(m, dtype=np.int32)
        rho_vec = np.full(m, settings.rho)
    rho_vec = np.asarray(rho_vec, dtype=np.float64)
    return rho_vec, 1.0 / rho_vec, constr_type


def p_full_pattern(p_upper: sp.csc_matrix):
    """Full symmetric P as (row, col, upper value index), upper entries first per pair
    (``solver.py:186-204``)."""
    rows, cols = csc_rows_cols(p_upper)
    ptr = np.arange(rows.size, dtype=np.int32)
    off = rows != cols
    # interleave: each upper entry, then its mirror when off-diagonal
    rep = np.where(off, 2, 1)
    idx = np.repeat(np.arange(rows.size), rep)
    first = np.ones(idx.size, dtype=bool)
    first[1:] = idx[1:] != idx[:-1]
    out_rows = np.where(first, rows[idx], cols[idx]).astype(np.int32)
    out_cols = np.where(first, cols[idx], rows[idx]).astype(np.int32)
    return out_rows, out_cols, ptr[idx]


def _col_norm_inf(matrix: sp.csc_matrix) -> np.ndarray:
    out = np.zeros(matrix.shape[1])
    for c in range(matrix.shape[1]):
        s, e = matrix.indptr[c], matrix.indptr[c + 1]
        if e > s:
            out[c] = np.max(np.abs(matrix.data[s:e]))
    return out


def _row_norm_inf(matrix: sp.csc_matrix) -> np.ndarray:
    out = np.zeros(matrix.shape[0])
    rows, _ = csc_rows_cols(matrix)
    np.maximum.at(out, rows, np.abs(matrix.data))
    return out


def _limit_vec(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, dtype=np.float64).copy()
    v[v < OSQP_MIN_SCALING] = 1.0
    v[v > OSQP_MAX_SCALING] = OSQP_MAX_SCALING
    return v


def _limit_scalar(v: float) -> float:
    if v < OSQP_MIN_SCALING:
        return 1.0
    return float(min(v, OSQP_MAX_SCALING))


def compute_scaling(p_upper: sp.csc_matrix, a_matrix: sp.csc_matrix, iterations: int, *, q=None) -> dict:
    """Modified Ruiz equilibration of (P, A[, q]) (``solver.py:262-343``)."""
    n = p_upper.shape[0]
    m = a_matrix.shape[0]
    d = np.ones(n)
    e = np.ones(m)
    c = 1.0
    p_rows, p_cols = csc_rows_cols(p_upper)
    a_rows, a_cols = csc_rows_cols(a_matrix)
    if iterations > 0:
        p_s = p_upper.copy().astype(np.float64)
        a_s = a_matrix.copy().astype(np.float64)
        q_s = None
        if q is not None:
            q_s = np.asarray(q, dtype=np.float64)
            if q_s.ndim == 2 and q_s.shape[0] == 1:
                q_s = q_s[0]
            if q_s.shape != (n,):
                raise ValueError("scaling_q must have shape (n,) or (1, n)")
            q_s = q_s.copy()
        for _ in range(iterations):
            d_t = _limit_vec(np.maximum(_col_norm_inf(p_s), _col_norm_inf(a_s)))
            e_t = _limit_vec(_row_norm_inf(a_s))
            d_t = 1.0 / np.sqrt(d_t)
            e_t = 1.0 / np.sqrt(e_t)
            p_s.data *= d_t[p_rows] * d_t[p_cols]
            a_s.data *= e_t[a_rows] * d_t[a_cols]
            if q_s is not None:
                q_s *= d_t
            d *= d_t
            e *= e_t
            p_cost = float(np.sum(_col_norm_inf(p_s)) / n)
            q_cost = 0.0 if q_s is None else float(np.max(np.abs(q_s)))
            q_cost = _limit_scalar(q_cost)
            gamma = 1.0 / _limit_scalar(max(p_cost, q_cost))
            p_s.data *= gamma
            if q_s is not None:
                q_s *= gamma
            c *= gamma
    return {
        "c": float(c),
        "cinv": float(1.0 / c),
        "d": d,
        "dinv": 1.0 / d,
        "e": e,
        "einv": 1.0 / e,
        "p_scale": c * d[p_rows] * d[p_cols],
        "a_scale": e[a_rows] * d[a_cols],
        "q_scale": c * d,
        "y_scale": c / e,
        "y_unscale": e / c,
    }


def build_osqp_plan(p_matrix, a_matrix, l=None, u=None, settings: OSQPSettings | None = None,
                    derivatives: bool = False, scaling_q=None, *, p=None, ordering: str = "amd") -> FixedOSQPPlan:
    """All fixed-pattern OSQP setup on the CPU (``solver.py:574-662``).

    ``p``/``ordering`` are forwarded to :func:`build_qdldl_plan` for the KKT.
    """
    settings = OSQPSettings() if settings is None else settings
    if settings.scaling < 0:
        raise ValueError("settings.scaling must be nonnegative")
    if settings.adaptive_rho:
        raise NotImplementedError("fixed-pattern OSQP expects adaptive_rho=False")
    if settings.polishing:
        raise NotImplementedError("fixed-pattern OSQP expects polishing=False")
    p_upper = upper_with_diagonal(p_matrix)
    a_csc = sp.csc_matrix(a_matrix).copy()
    a_csc.sum_duplicates()
    a_csc.sort_indices()
    if p_upper.shape[0] != p_upper.shape[1]:
        raise ValueError("P must be square")
    if a_csc.shape[1] != p_upper.shape[0]:
        raise ValueError("A column count must match P dimension")
    sc = compute_scaling(p_upper, a_csc, settings.scaling, q=scaling_q)
    scaled_l = None if l is None else np.asarray(l, dtype=np.float64) * sc["e"]
    scaled_u = None if u is None else np.asarray(u, dtype=np.float64) * sc["e"]
    rho_vec, rho_inv_vec, constr_type = rho_vectors(settings, a_csc.shape[0], scaled_l, scaled_u)
    p_kkt = p_upper.copy().astype(np.float64)
    p_kkt.data = p_kkt.data * sc["p_scale"]
    a_kkt = a_csc.copy().astype(np.float64)
    a_kkt.data = a_kkt.data * sc["a_scale"]
    kkt, p_to_kkt, p_diag_sigma, a_to_kkt, rho_to_kkt = build_kkt_and_maps(p_kkt, a_kkt, settings, rho_inv_vec)
    qplan = build_qdldl_plan(kkt, upper=True, p=p, ordering=ordering)
    pf_rows, pf_cols, pf_map = p_full_pattern(p_upper)
    derivative_plan = None
    if derivatives:
        from .derivatives import build_derivative_plan

        derivative_plan = build_derivative_plan(p_upper, a_csc, l, u, ordering=ordering)
    a_rows, a_cols = csc_rows_cols(a_csc)
    return FixedOSQPPlan(
        n=p_upper.shape[0],
        m=a_csc.shape[0],
        p_upper=p_upper,
        a_matrix=a_csc,
        kkt_upper=kkt,
        qdldl_plan=qplan,
        p_to_kkt=p_to_kkt,
        p_diag_sigma=p_diag_sigma,
        a_to_kkt=a_to_kkt,
        rho_to_kkt=rho_to_kkt,
        a_rows=a_rows,
        a_cols=a_cols,
        p_full_rows=pf_rows,
        p_full_cols=pf_cols,
        p_full_value_map=pf_map,
        rho_vec=rho_vec,
        rho_inv_vec=rho_inv_vec,
        constr_type=constr_type,
        scaling_c=sc["c"],
        scaling_cinv=sc["cinv"],
        scaling_d=sc["d"],
        scaling_dinv=sc["dinv"],
        scaling_e=sc["e"],
        scaling_einv=sc["einv"],
        p_scale=sc["p_scale"],
        a_scale=sc["a_scale"],
        q_scale=sc["q_scale"],
        y_scale=sc["y_scale"],
        y_unscale=sc["y_unscale"],
        derivative_plan=derivative_plan,
        settings=settings,
    )


__all__ = [
    "OSQP_INFTY",
    "bounds_type",
    "build_kkt_and_maps",
    "build_osqp_plan",
    "compute_scaling",
    "csc_rows_cols",
    "p_full_pattern",
    "rho_vectors",
    "upper_with_diagonal",
]
