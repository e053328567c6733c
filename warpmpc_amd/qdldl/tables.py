"""Ragged (unpadded) program tables derived from a :class:`QDLDLPlan`.

The plan stores the reference's padded ``(rows, max_width)`` tables
(``warpmpc/jax_qdldl/core.py:228-256``).  Every mask is a prefix mask, so the
padding can be dropped without changing any order; the result is a set of
CSR-like int32 arrays that the HIP library and the CPU oracle consume.
"""

from __future__ import annotations

import numpy as np

from .types import QDLDLPlan


def _ragged(values: np.ndarray, mask: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    counts = mask.sum(axis=1).astype(np.int64)
    if mask.size and not np.array_equal(mask, np.arange(mask.shape[1])[None, :] < counts[:, None]):
        raise ValueError("program mask is not a prefix mask")
    ptr = np.zeros(mask.shape[0] + 1, dtype=np.int32)
    np.cumsum(counts, out=ptr[1:])
    return ptr, np.ascontiguousarray(values[mask], dtype=np.int32)


def ragged_program(plan: QDLDLPlan) -> dict[str, np.ndarray]:
    init_ptr, init_rows = _ragged(plan.init_rows, plan.init_mask)
    _, init_pos = _ragged(plan.init_pos, plan.init_mask)
    proc_ptr, proc_cols = _ragged(plan.process_cols, plan.process_mask)
    _, proc_lidx = _ragged(plan.process_lidx, plan.process_mask)
    _, proc_off = _ragged(plan.process_offsets, plan.process_mask)
    as32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
    return {
        "perm_to_orig": as32(plan.perm_to_orig),
        "diag_pos": as32(plan.diag_pos),
        "init_ptr": init_ptr,
        "init_rows": init_rows,
        "init_pos": init_pos,
        "proc_ptr": proc_ptr,
        "proc_cols": proc_cols,
        "proc_lidx": proc_lidx,
        "proc_off": proc_off,
        "lp": as32(plan.lp),
        "li": as32(plan.li),
        "p": as32(plan.p),
        "pinv": as32(plan.pinv),
    }


__all__ = ["ragged_program"]
