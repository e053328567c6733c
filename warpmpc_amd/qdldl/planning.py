"""Host-side symbolic planning of a fixed-pattern batched LDLᵀ (QDLDL).

Restates the reference planner (``warpmpc/jax_qdldl/core.py:38-311``):

* upper-triangular CSC normalisation (``_as_upper_csc``, ``core.py:38-47``),
* symmetric permutation of the upper triangle (``_symperm_upper``,
  ``core.py:57-102``; AMD's ``perm.c`` semantics, entries appended in the order
  the original columns are traversed),
* the QDLDL elimination tree and column counts (``_qdldl_etree``,
  ``core.py:105-131``),
* the up-looking numeric "program" of every row (``_build_numeric_program``,
  ``core.py:147-256``).

The fill-reducing permutation comes from :func:`ordering.amd_order` (own AMD)
unless the caller supplies ``p`` explicitly.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .ordering import amd_order, etree_postorder, natural_order
from .types import QDLDLPlan

QDLDL_UNKNOWN = -1


def as_upper_csc(matrix, upper: bool) -> sp.csc_matrix:
    """Upper triangle in canonical CSC (``core.py:38-47``)."""
    matrix = sp.csc_matrix(matrix)
    out = matrix.copy() if upper else sp.triu(matrix, format="csc")
    out.sum_duplicates()
    out.sort_indices()
    return out


def inverse_permutation(p: np.ndarray) -> np.ndarray:
    pinv = np.empty_like(p)
    pinv[p] = np.arange(p.size, dtype=p.dtype)
    return pinv


def symperm_upper(matrix: sp.csc_matrix, pinv: np.ndarray):
    """Upper CSC of ``P A Pᵀ`` plus the value map ``orig_to_perm``.

    Same result as ``core.py:57-102``: an entry ``(i, j)`` (``i <= j``) moves to
    column ``max(pinv[i], pinv[j])``; inside a destination column entries keep the
    order in which the original columns (then pointers) are visited.
    """
    n = matrix.shape[0]
    ap = np.asarray(matrix.indptr, dtype=np.int64)
    ai = np.asarray(matrix.indices, dtype=np.int64)
    cols = np.repeat(np.arange(n, dtype=np.int64), np.diff(ap))
    if np.any(ai > cols):
        raise ValueError("input matrix contains lower-triangular entries")
    i2 = pinv[ai].astype(np.int64)
    j2 = pinv[cols].astype(np.int64)
    dest_col = np.maximum(i2, j2)
    dest_row = np.minimum(i2, j2)
    # stable sort by destination column keeps the traversal order inside a column
    order = np.argsort(dest_col, kind="stable")
    counts = np.bincount(dest_col, minlength=n)
    cp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=cp[1:])
    ci = dest_row[order].astype(np.int32)
    orig_to_perm = np.empty(ai.size, dtype=np.int32)
    orig_to_perm[order] = np.arange(ai.size, dtype=np.int32)
    return cp.astype(np.int32), ci, orig_to_perm


def qdldl_etree(n: int, ap: np.ndarray, ai: np.ndarray):
    """Elimination tree and L column counts of an upper CSC pattern (``core.py:105-131``)."""
    work = np.zeros(n, dtype=np.int64)
    lnz = np.zeros(n, dtype=np.int32)
    etree = np.full(n, QDLDL_UNKNOWN, dtype=np.int32)
    if np.any(np.diff(ap) == 0):
        col = int(np.flatnonzero(np.diff(ap) == 0)[0])
        raise ValueError(f"column {col} is empty; QDLDL requires a nonempty column")
    ap_l = ap.tolist()
    ai_l = ai.tolist()
    et = [QDLDL_UNKNOWN] * n
    cnt = [0] * n
    wk = [0] * n
    for j in range(n):
        wk[j] = j
        for ptr in range(ap_l[j], ap_l[j + 1]):
            i = ai_l[ptr]
            if i > j:
                raise ValueError("matrix is not upper triangular after permutation")
            while wk[i] != j:
                if et[i] == QDLDL_UNKNOWN:
                    et[i] = j
                cnt[i] += 1
                wk[i] = j
                i = et[i]
    etree[:] = et
    lnz[:] = cnt
    del work
    return lnz, etree, int(np.sum(lnz, dtype=np.int64))


def _pad(rows: list[list[int]], dtype=np.int32):
    width = max((len(r) for r in rows), default=0)
    out = np.zeros((len(rows), width), dtype=dtype)
    mask = np.zeros((len(rows), width), dtype=bool)
    for i, r in enumerate(rows):
        if r:
            out[i, : len(r)] = r
            mask[i, : len(r)] = True
    return out, mask


def row_reach(k: int, ap, ai, etree, marker: list[int]) -> list[int]:
    """Columns of row ``k`` of L in QDLDL's processing order.

    For every off-diagonal entry of column ``k`` the etree path up to ``k`` is
    walked; each path is pushed so that, read from the end, ancestors come after
    descendants.  The returned list is already in processing order (QDLDL walks
    ``yIdx`` from its end), matching ``process_cols`` of ``core.py:196-226``.
    """
    y_idx: list[int] = []
    for ptr in range(ap[k], ap[k + 1]):
        b = ai[ptr]
        if b == k or marker[b] == k:
            continue
        path = [b]
        marker[b] = k
        nxt = etree[b]
        while nxt != QDLDL_UNKNOWN and nxt < k and marker[nxt] != k:
            marker[nxt] = k
            path.append(nxt)
            nxt = etree[nxt]
        y_idx.extend(reversed(path))
    y_idx.reverse()
    return y_idx


def build_numeric_program(n: int, ap: np.ndarray, ai: np.ndarray, lnz: np.ndarray, etree: np.ndarray):
    """Per-row programs of the up-looking factorization (``core.py:147-256``).

    Row ``k``:
      * ``init_rows/init_pos``: off-diagonal entries of permuted column ``k``;
      * ``process_cols``: the reach of row ``k`` in processing order, with
        ``process_lidx`` (slot of ``L[k, c]``) and ``process_offsets`` (how many
        entries column ``c`` already holds, i.e. its rows ``< k``).
    Column ``c``: ``col_rows/col_lidx`` list the entries of L column ``c``.
    """
    lp = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(lnz, out=lp[1:])
    nnz_l = int(lp[-1])
    li = np.empty(nnz_l, dtype=np.int32)
    next_slot = lp[:-1].astype(np.int64).copy()
    diag_pos = np.full(n, -1, dtype=np.int32)
    ap_l = ap.tolist()
    ai_l = ai.tolist()
    et = etree.tolist()
    marker = [-1] * n
    init_rows, init_pos, p_cols, p_lidx, p_off = [], [], [], [], []
    for k in range(n):
        rows_k, pos_k = [], []
        for ptr in range(ap_l[k], ap_l[k + 1]):
            r = ai_l[ptr]
            if r == k:
                diag_pos[k] = ptr
            else:
                rows_k.append(r)
                pos_k.append(ptr)
        if diag_pos[k] < 0:
            raise ValueError(f"column {k} has no diagonal entry")
        init_rows.append(rows_k)
        init_pos.append(pos_k)
        cols = row_reach(k, ap_l, ai_l, et, marker) if k > 0 else []
        lidx, offs = [], []
        for c in cols:
            dest = int(next_slot[c])
            lidx.append(dest)
            offs.append(dest - int(lp[c]))
            li[dest] = k
            next_slot[c] += 1
        p_cols.append(cols)
        p_lidx.append(lidx)
        p_off.append(offs)
    if not np.array_equal(next_slot, lp[1:]):
        raise AssertionError("symbolic program did not fill L")

    init_rows_a, init_mask = _pad(init_rows)
    init_pos_a, _ = _pad(init_pos)
    process_cols, process_mask = _pad(p_cols)
    process_lidx, _ = _pad(p_lidx)
    process_offsets, _ = _pad(p_off)
    col_rows_l = [li[lp[c] : lp[c + 1]].tolist() for c in range(n)]
    col_lidx_l = [list(range(int(lp[c]), int(lp[c + 1]))) for c in range(n)]
    col_rows, col_mask = _pad(col_rows_l)
    col_lidx, _ = _pad(col_lidx_l)
    return {
        "lp": lp,
        "li": li,
        "diag_pos": diag_pos,
        "init_rows": init_rows_a,
        "init_pos": init_pos_a,
        "init_mask": init_mask,
        "process_cols": process_cols,
        "process_lidx": process_lidx,
        "process_offsets": process_offsets,
        "process_mask": process_mask,
        "col_rows": col_rows,
        "col_lidx": col_lidx,
        "col_mask": col_mask,
    }


def build_qdldl_plan(
    matrix,
    upper: bool = False,
    *,
    p: np.ndarray | None = None,
    ordering: str = "amd",
) -> QDLDLPlan:
    """Plan a fixed-pattern batched LDLᵀ of a symmetric quasi-definite matrix.

    Same call as ``build_qdldl_plan(matrix, upper=False)`` (``core.py:259-311``).
    ``p`` supplies an external permutation (e.g. the one the reference's qdldl
    AMD produced) so L/D can be compared entry by entry; otherwise ``ordering``
    selects ``"amd"`` (own approximate minimum degree) or ``"natural"``.
    """
    a_upper = as_upper_csc(matrix, upper=upper)
    n = a_upper.shape[0]
    if a_upper.shape[0] != a_upper.shape[1]:
        raise ValueError("QDLDL matrix must be square")
    if p is not None:
        p = np.asarray(p, dtype=np.int32)
        if p.shape != (n,) or not np.array_equal(np.sort(p), np.arange(n)):
            raise ValueError("p must be a permutation of range(n)")
        ordering = "external"
    elif ordering == "amd":
        p = amd_order(n, a_upper.indptr, a_upper.indices)
        # relabel by a largest-subtree-first postorder of the elimination tree
        # (same fill, short column lifetimes; see ordering.etree_postorder)
        ap0, ai0, _ = symperm_upper(a_upper, inverse_permutation(p))
        _, parent0, _ = qdldl_etree(n, ap0, ai0)
        p = np.ascontiguousarray(p[etree_postorder(parent0)], dtype=np.int32)
    elif ordering == "natural":
        p = natural_order(n)
    else:
        raise ValueError(f"unknown ordering {ordering!r}")
    pinv = inverse_permutation(p)
    aperm_p, aperm_i, orig_to_perm = symperm_upper(a_upper, pinv)
    perm_to_orig = np.empty_like(orig_to_perm)
    perm_to_orig[orig_to_perm] = np.arange(orig_to_perm.size, dtype=np.int32)
    lnz, etree, nnz_l = qdldl_etree(n, aperm_p, aperm_i)
    program = build_numeric_program(n, aperm_p, aperm_i, lnz, etree)
    return QDLDLPlan(
        n=n,
        nnz_a=int(a_upper.indptr[-1]),
        nnz_l=nnz_l,
        a_indptr=np.asarray(a_upper.indptr, dtype=np.int32),
        a_indices=np.asarray(a_upper.indices, dtype=np.int32),
        a_data=np.asarray(a_upper.data),
        aperm_indptr=aperm_p,
        aperm_indices=aperm_i,
        perm_to_orig=perm_to_orig,
        orig_to_perm=orig_to_perm,
        p=p,
        pinv=pinv,
        etree=etree,
        lnz=lnz,
        ordering=ordering,
        **program,
    )


__all__ = [
    "QDLDL_UNKNOWN",
    "as_upper_csc",
    "build_numeric_program",
    "build_qdldl_plan",
    "inverse_permutation",
    "qdldl_etree",
    "row_reach",
    "symperm_upper",
]
