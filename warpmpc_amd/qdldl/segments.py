"""Segmentation of the row/column scans (the ``segmented`` option).

Restates ``warpmpc/jax_qdldl/core.py:320-485``:

* ``fixed`` segments: ``budget`` nearly equal contiguous pieces;
* ``greedy`` segments: repeatedly apply the split with the largest saving of
  padded work ``(stop - start) * max(counts[start:stop])``;
* ``optimal`` segments: exact minimum of the total padded work over all
  contiguous splits into ``min(budget, n)`` pieces, by dynamic programming.

The HIP kernels walk ragged (unpadded) tables, so segments do not change what
they run; ``compile_qdldl_variant`` uses these routines to validate the
``segmented`` options as the reference does.
"""

from __future__ import annotations

import numpy as np


def _counts(mask: np.ndarray) -> np.ndarray:
    return np.asarray(np.sum(mask, axis=1), dtype=np.int64)


def _check(n: int, budget: int) -> None:
    if budget <= 0:
        raise ValueError("segment_budget must be positive")


def make_fixed_segments(mask: np.ndarray, segment_budget: int):
    n = int(mask.shape[0])
    if n == 0:
        return []
    _check(n, segment_budget)
    edges = np.linspace(0, n, min(int(segment_budget), n) + 1, dtype=np.int64)
    return [(int(a), np.arange(a, b, dtype=np.int32)) for a, b in zip(edges[:-1], edges[1:])]


def _best_split(counts: np.ndarray, start: int, stop: int):
    length = stop - start
    if length <= 1:
        return None, 0
    seg = counts[start:stop]
    whole = length * int(seg.max(initial=0))
    left = np.maximum.accumulate(seg[:-1])
    right = np.maximum.accumulate(seg[::-1][:-1])[::-1]
    at = np.arange(1, length, dtype=np.int64)
    saving = whole - at * left - (length - at) * right
    i = int(np.argmax(saving))
    if saving[i] <= 0:
        return None, 0
    return start + i + 1, int(saving[i])


def make_greedy_segments(mask: np.ndarray, segment_budget: int):
    n = int(mask.shape[0])
    if n == 0:
        return []
    _check(n, segment_budget)
    counts = _counts(mask)
    pieces = [(0, n)]
    target = min(int(segment_budget), n)
    while len(pieces) < target:
        best = (-1, None, 0)
        for i, (a, b) in enumerate(pieces):
            split, saving = _best_split(counts, a, b)
            if saving > best[2]:
                best = (i, split, saving)
        i, split, saving = best
        if split is None or saving <= 0:
            break
        a, b = pieces[i]
        pieces[i : i + 1] = [(a, split), (split, b)]
    return [(a, np.arange(a, b, dtype=np.int32)) for a, b in pieces]


def make_optimal_segments(mask: np.ndarray, segment_budget: int):
    """Exact DP ``dp[k][j] = min_i dp[k-1][i] + (j - i) * max(counts[i:j])``.

    For every end column the candidate starts form runs with a common running
    maximum (a monotone stack of backward record maxima); per run only
    ``min_i dp[k-1][i] - i * run_max`` matters, so each end scans only the stack.
    """
    n = int(mask.shape[0])
    if n == 0:
        return []
    _check(n, segment_budget)
    target = min(int(segment_budget), n)
    counts = _counts(mask)
    idx = np.arange(n, dtype=np.float64)
    prev = np.full(n + 1, np.inf)
    prev[0] = 0.0
    back = np.full((target + 1, n + 1), -1, dtype=np.int64)
    for k in range(1, target + 1):
        cur = np.full(n + 1, np.inf)
        stack: list[tuple[int, int, float, int]] = []  # (first start, run max, best value, best start)
        for t in range(n):
            cmax = int(counts[t])
            first = t
            while stack and stack[-1][1] <= cmax:
                first = stack.pop()[0]
            vals = prev[first : t + 1] - idx[first : t + 1] * cmax
            j = int(np.argmin(vals))
            stack.append((first, cmax, float(vals[j]), first + j))
            end = t + 1
            if end < k:
                continue
            best, arg = np.inf, -1
            for _, rmax, rbest, rstart in stack:
                cand = rbest + end * rmax
                if cand < best:
                    best, arg = cand, rstart
            cur[end] = best
            back[k, end] = arg
        prev = cur
    pieces = []
    end = n
    for k in range(target, 0, -1):
        start = int(back[k, end])
        if start < 0:
            raise AssertionError("segmentation backtracking failed")
        pieces.append((start, end))
        end = start
    if end != 0:
        raise AssertionError("segmentation does not cover every column")
    pieces.reverse()
    return [(a, np.arange(a, b, dtype=np.int32)) for a, b in pieces]


def make_segments(mask: np.ndarray, *, segment_budget: int, segment_strategy: str = "optimal"):
    if segment_strategy == "fixed":
        return make_fixed_segments(mask, segment_budget)
    if segment_strategy == "greedy":
        return make_greedy_segments(mask, segment_budget)
    if segment_strategy == "optimal":
        return make_optimal_segments(mask, segment_budget)
    raise ValueError(f"unknown segment_strategy {segment_strategy!r}")


__all__ = [
    "make_fixed_segments",
    "make_greedy_segments",
    "make_optimal_segments",
    "make_segments",
]
