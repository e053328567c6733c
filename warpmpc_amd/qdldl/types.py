"""Plan and compiled-solver containers for the fixed-pattern batched QDLDL.

Field names and meanings follow ``warpmpc/jax_qdldl/types.py:11-70`` so that a
caller written against the reference's ``QDLDLPlan`` / ``CompiledQDLDL`` reads
the same attributes here.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable

import numpy as np


@dataclass(frozen=True)
class QDLDLPlan:
    """Every pattern-dependent array of one fixed KKT sparsity pattern.

    Mirrors ``QDLDLPlan`` (``warpmpc/jax_qdldl/types.py:11-53``):

    * ``a_indptr/a_indices/a_data``: the unpermuted upper-triangular CSC input
      (value arrays passed to ``factor`` follow this order).
    * ``p``/``pinv``: fill-reducing permutation (permuted index ``i`` is original
      index ``p[i]``).
    * ``aperm_*``, ``perm_to_orig``, ``orig_to_perm``: upper CSC of ``P A Pᵀ`` and
      the value maps between the two orders.
    * ``etree``, ``lnz``, ``lp``, ``li``: elimination tree and L pattern (CSC,
      strictly lower, unit diagonal implicit).
    * ``init_*``, ``process_*``, ``col_*``: padded per-row / per-column programs of
      the up-looking factorization (see ``planning.build_numeric_program``).
    """

    n: int
    nnz_a: int
    nnz_l: int
    a_indptr: np.ndarray
    a_indices: np.ndarray
    a_data: np.ndarray
    aperm_indptr: np.ndarray
    aperm_indices: np.ndarray
    perm_to_orig: np.ndarray
    orig_to_perm: np.ndarray
    p: np.ndarray
    pinv: np.ndarray
    etree: np.ndarray
    lnz: np.ndarray
    lp: np.ndarray
    li: np.ndarray
    diag_pos: np.ndarray
    init_rows: np.ndarray
    init_pos: np.ndarray
    init_mask: np.ndarray
    process_cols: np.ndarray
    process_lidx: np.ndarray
    process_offsets: np.ndarray
    process_mask: np.ndarray
    col_rows: np.ndarray
    col_lidx: np.ndarray
    col_mask: np.ndarray
    ordering: str = "amd"

    @property
    def max_init(self) -> int:
        return int(self.init_rows.shape[1])

    @property
    def max_row_nnz(self) -> int:
        return int(self.process_cols.shape[1])

    @property
    def max_col_nnz(self) -> int:
        return int(self.col_rows.shape[1])


@dataclass(frozen=True)
class CompiledQDLDL:
    """Fixed-pattern factor / solve callables bound to one plan.

    Same contract as ``CompiledQDLDL`` (``warpmpc/jax_qdldl/types.py:56-70``) in
    the symbolic layout the Warp backend uses (``warp_backend.py:1410-1429``):

    * ``factor(a_values[nnz_a, B]) -> (lx[nnz_l, B], d[n, B], dinv[n, B])``
    * ``solve(lx, dinv, rhs[n, B]) -> x[n, B]``
    * ``factor_and_solve(a_values, rhs) -> (x, lx, d)``
    """

    plan: QDLDLPlan
    dtype: np.dtype
    factor: Callable
    solve: Callable
    factor_and_solve: Callable
    variant: str = "baseline"
    values_layout: str = "original_symbolic"
    rhs_layout: str = "symbolic"
    backend: str = "hip"
    factor_backend: str = "hip"
    solve_backend: str = "hip"
    extras: dict = field(default_factory=dict, compare=False, repr=False)


__all__ = ["CompiledQDLDL", "QDLDLPlan"]
