"""Pattern data of OSQP's adjoint (derivative) linear system.

Restates ``_build_derivative_plan`` (``warpmpc/jax_osqp/solver.py:413-571``) and
its helpers ``_full_symmetric_csc_pattern`` / ``_selected_a_csc_pattern``
(``solver.py:346-410``).  With the constraint rows split into lower-bounded
inequalities (``l > -1e26``, ``l < u``), upper-bounded inequalities
(``u < 1e26``, ``l < u``) and equalities (``l == u``), and

    G = [-A_lower; A_upper],  h = [-l_lower; u_upper],  Aeq = A_eq,
    half = n + n_ineq + n_eq,

the adjoint matrix is the symmetric ``2 half x 2 half`` matrix

    M = [[(1 + eps) I, U], [Uᵀ, -eps I]],
    U = [[P, Gᵀ diag(lambda), Aeqᵀ], [G, diag(G x - h), 0], [Aeq, 0, 0]],

stored as its upper triangle; ``M0`` is the same with ``eps = 0`` on both
diagonals (the matrix the iterative refinement converges to).  The block
layout, the entry order inside each block and ``eps = 1e-6`` follow the
reference so that ``adjoint_upper``'s structure matches it.

The HIP library derives its per-slot value sources and gather tables from
these fields (``wmpc_adjoint_plan_create``).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from ..qdldl.planning import build_qdldl_plan
from .types import OSQPDerivativePlan

OSQP_INFTY = 1e30
OSQP_MIN_SCALING = 1e-4
ADJOINT_EPS = 1e-6


def _csc_rows_cols(matrix: sp.csc_matrix):
    cols = np.repeat(np.arange(matrix.shape[1], dtype=np.int32), np.diff(matrix.indptr))
    return np.asarray(matrix.indices, dtype=np.int32), cols


def full_symmetric_pattern(p_upper: sp.csc_matrix):
    """Full symmetric pattern of P in CSC order and, per entry, the index of the
    upper-triangular entry that holds its value (``solver.py:346-374``)."""
    rows, cols = _csc_rows_cols(p_upper)
    idx = np.arange(rows.size, dtype=np.int64)
    off = rows != cols
    fr = np.concatenate([rows, cols[off]])
    fc = np.concatenate([cols, rows[off]])
    fv = np.concatenate([idx, idx[off]])
    order = np.lexsort((fr, fc))
    return fr[order].astype(np.int32), fc[order].astype(np.int32), fv[order].astype(np.int32)


def selected_rows_pattern(a_csc: sp.csc_matrix, selected: np.ndarray, row_offset: int = 0):
    """Entries of the rows ``selected`` of A, renumbered ``row_offset + position``,
    in CSC order, with the index of each A value (``solver.py:377-410``)."""
    selected = np.asarray(selected, dtype=np.int64)
    pos = np.full(a_csc.shape[0], -1, dtype=np.int64)
    pos[selected] = np.arange(selected.size)
    rows, cols = _csc_rows_cols(a_csc)
    keep = pos[rows] >= 0
    out_r = row_offset + pos[rows[keep]]
    out_c = cols[keep]
    vidx = np.flatnonzero(keep)
    order = np.lexsort((out_r, out_c))
    return out_r[order].astype(np.int32), out_c[order].astype(np.int32), vidx[order].astype(np.int32)


def build_derivative_plan(p_upper: sp.csc_matrix, a_csc: sp.csc_matrix, l, u, *, ordering: str = "amd",
                          p=None) -> OSQPDerivativePlan | None:
    """OSQP's adjoint-system plan (``solver.py:413-571``); None without bounds."""
    if l is None or u is None:
        return None
    n = p_upper.shape[0]
    m = a_csc.shape[0]
    l = np.asarray(l, dtype=float)
    u = np.asarray(u, dtype=float)
    infval = OSQP_INFTY * OSQP_MIN_SCALING
    is_ineq = l < u
    lower = np.flatnonzero(is_ineq & (l > -infval)).astype(np.int32)
    upper = np.flatnonzero(is_ineq & (u < infval)).astype(np.int32)
    eq = np.flatnonzero(~is_ineq).astype(np.int32)
    n_l, n_u, n_eq = lower.size, upper.size, eq.size
    n_ineq = n_l + n_u
    half = n + n_ineq + n_eq
    dim = 2 * half

    p_rows, p_cols = _csc_rows_cols(p_upper)
    pf_rows, pf_cols, pf_map = full_symmetric_pattern(p_upper)
    lr, lc, lm = selected_rows_pattern(a_csc, lower, 0)
    ur, uc, um = selected_rows_pattern(a_csc, upper, n_l)
    g_rows = np.concatenate([lr, ur])
    g_cols = np.concatenate([lc, uc])
    g_map = np.concatenate([lm, um])
    g_sign = np.concatenate([-np.ones(lr.size), np.ones(ur.size)])
    order = np.lexsort((g_rows, g_cols))
    g_rows, g_cols, g_map, g_sign = g_rows[order], g_cols[order], g_map[order], g_sign[order]
    e_rows, e_cols, e_map = selected_rows_pattern(a_csc, eq, 0)

    # entries of the upper triangle, block by block in the reference's order
    blocks = []  # (tag, rows, cols)
    ar = np.arange(half, dtype=np.int64)
    blocks.append(("diag_first", ar, ar))
    blocks.append(("p", pf_rows.astype(np.int64), half + pf_cols.astype(np.int64)))
    blocks.append(("g", n + g_rows.astype(np.int64), half + g_cols.astype(np.int64)))
    blocks.append(("aeq", n + n_ineq + e_rows.astype(np.int64), half + e_cols.astype(np.int64)))
    blocks.append(("glambda", g_cols.astype(np.int64), half + n + g_rows.astype(np.int64)))
    ai = np.arange(n_ineq, dtype=np.int64)
    blocks.append(("slack", n + ai, half + n + ai))
    blocks.append(("aeq_t", e_cols.astype(np.int64), half + n + n_ineq + e_rows.astype(np.int64)))
    blocks.append(("diag_second", half + ar, half + ar))
    all_r = np.concatenate([b[1] for b in blocks])
    all_c = np.concatenate([b[2] for b in blocks])
    if np.any(all_r > all_c):
        raise AssertionError("adjoint entries must lie in the upper triangle")
    order = np.lexsort((all_r, all_c))
    pos = np.empty(order.size, dtype=np.int64)
    pos[order] = np.arange(order.size)
    if np.unique(all_c * dim + all_r).size != all_r.size:
        raise AssertionError("adjoint blocks overlap")
    counts = np.bincount(all_c, minlength=dim)
    indptr = np.zeros(dim + 1, dtype=np.int32)
    np.cumsum(counts, out=indptr[1:])
    rows_sorted = all_r[order].astype(np.int32)
    positions = {}
    start = 0
    for tag, rr, _ in blocks:
        positions[tag] = pos[start : start + rr.size].astype(np.int32)
        start += rr.size

    eps = ADJOINT_EPS
    data = np.zeros(order.size)
    data[positions["diag_first"]] = 1.0 + eps
    data[positions["diag_second"]] = -eps
    data[positions["p"]] = np.asarray(p_upper.data, dtype=float)[pf_map]
    data[positions["g"]] = np.asarray(a_csc.data, dtype=float)[g_map] * g_sign
    data[positions["aeq"]] = np.asarray(a_csc.data, dtype=float)[e_map]
    data[positions["glambda"]] = 0.0
    data[positions["slack"]] = -1.0
    data[positions["aeq_t"]] = np.asarray(a_csc.data, dtype=float)[e_map]
    adjoint_upper = sp.csc_matrix((data, rows_sorted, indptr), shape=(dim, dim))
    qplan = build_qdldl_plan(adjoint_upper, upper=True, p=p, ordering=ordering)
    return OSQPDerivativePlan(
        n=n,
        m=m,
        n_ineq_l=n_l,
        n_ineq_u=n_u,
        n_eq=n_eq,
        adjoint_dim=dim,
        adjoint_upper=adjoint_upper,
        qdldl_plan=qplan,
        lower_indices=lower,
        upper_indices=upper,
        eq_indices=eq,
        p_upper_rows=p_rows,
        p_upper_cols=p_cols,
        p_full_rows=pf_rows,
        p_full_cols=pf_cols,
        p_full_value_map=pf_map,
        g_rows=g_rows.astype(np.int32),
        g_cols=g_cols.astype(np.int32),
        g_a_value_map=g_map.astype(np.int32),
        g_sign=g_sign,
        aeq_rows=e_rows,
        aeq_cols=e_cols,
        aeq_a_value_map=e_map,
        diag_first_pos=positions["diag_first"],
        diag_second_pos=positions["diag_second"],
        p_block_pos=positions["p"],
        g_block_pos=positions["g"],
        aeq_block_pos=positions["aeq"],
        glambda_block_pos=positions["glambda"],
        slack_diag_pos=positions["slack"],
        aeq_t_block_pos=positions["aeq_t"],
        eps=eps,
    )


__all__ = [
    "ADJOINT_EPS",
    "build_derivative_plan",
    "full_symmetric_pattern",
    "selected_rows_pattern",
]
