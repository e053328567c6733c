"""Level sets and the level schedule of the triangular solves.

Restates ``_make_solve_levels`` (``warpmpc/jax_qdldl/core.py:488-510``),
``_pack_levels`` and ``_make_level_schedule`` (``warp_backend.py:1098-1126``)
on a plan's L pattern.  The level-scheduled HIP solve
(``wmpc_qdldl_solve_levels_f32``) derives the same levels from the plan at
creation and applies the schedule's rule on the device: a level with more
than ``threshold`` columns is one parallel step, a run of levels of at most
``threshold`` columns one serial step.
"""

from __future__ import annotations

import numpy as np

from .types import QDLDLPlan


def make_solve_levels(plan: QDLDLPlan) -> list[np.ndarray]:
    """Columns of L by level: an entry (row, col) makes ``row`` depend on
    ``col``; level(row) = max over its columns of level(col) + 1."""
    n = int(plan.n)
    lp = np.asarray(plan.lp, dtype=np.int64)
    li = np.asarray(plan.li, dtype=np.int64)
    levels = np.zeros(n, dtype=np.int32)
    for col in range(n):
        rows = li[lp[col]:lp[col + 1]]
        if rows.size:
            levels[rows] = np.maximum(levels[rows], levels[col] + 1)
    top = int(levels.max(initial=0))
    return [np.flatnonzero(levels == lev).astype(np.int32) for lev in range(top + 1)]


def pack_levels(levels: list[np.ndarray]) -> tuple[np.ndarray, np.ndarray]:
    """(flat column list, level offsets)."""
    offsets = np.zeros(len(levels) + 1, dtype=np.int32)
    offsets[1:] = np.cumsum([len(lv) for lv in levels])
    flat = np.concatenate(levels).astype(np.int32) if levels else np.zeros(0, dtype=np.int32)
    return flat, offsets


def make_level_schedule(levels: list[np.ndarray], offsets: np.ndarray, *,
                        serial_width_threshold: int) -> tuple[tuple[str, int, int, int, int], ...]:
    """("serial", first level, end level, first offset, 0) for a maximal run of
    levels of at most ``serial_width_threshold`` columns, ("parallel", level,
    level + 1, offset, width) for each wider level."""
    out = []
    lev = 0
    while lev < len(levels):
        width = int(len(levels[lev]))
        if width <= serial_width_threshold:
            start = lev
            lev += 1
            while lev < len(levels) and int(len(levels[lev])) <= serial_width_threshold:
                lev += 1
            out.append(("serial", start, lev, int(offsets[start]), 0))
        else:
            out.append(("parallel", lev, lev + 1, int(offsets[lev]), width))
            lev += 1
    return tuple(out)


__all__ = ["make_level_schedule", "make_solve_levels", "pack_levels"]
