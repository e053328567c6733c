"""Segmentation known answers of tests/test_jax_qdldl_segments.py:56-88 (CPU only)."""

from __future__ import annotations

import itertools

import numpy as np

from warpmpc_amd.qdldl.segments import (
    make_fixed_segments,
    make_greedy_segments,
    make_level_schedule,
    make_optimal_segments,
    make_segments,
    make_solve_levels,
)


def _mask(counts):
    width = max(counts, default=0)
    mask = np.zeros((len(counts), width), dtype=bool)
    for i, c in enumerate(counts):
        mask[i, :c] = True
    return mask


def _cost(counts, segments):
    return sum((int(c[-1]) + 1 - int(c[0])) * int(np.max(counts[int(c[0]) : int(c[-1]) + 1], initial=0))
               for _, c in segments)


def _intervals(segments):
    return [(int(c[0]), int(c[-1]) + 1) for _, c in segments]


def _brute(counts, k):
    n = len(counts)
    best = None
    for splits in itertools.combinations(range(1, n), k - 1):
        starts = (0, *splits)
        stops = (*splits, n)
        cost = sum((b - a) * int(np.max(counts[a:b], initial=0)) for a, b in zip(starts, stops))
        best = cost if best is None else min(best, cost)
    return best


def test_strategy_selector_uses_segment_budget():
    counts = np.array([5, 1, 4, 1, 3])
    mask = _mask(counts.tolist())
    fixed = make_segments(mask, segment_budget=3, segment_strategy="fixed")
    greedy = make_segments(mask, segment_budget=3, segment_strategy="greedy")
    optimal = make_segments(mask, segment_budget=3, segment_strategy="optimal")
    assert _intervals(fixed) == _intervals(make_fixed_segments(mask, 3))
    assert _intervals(greedy) == _intervals(make_greedy_segments(mask, 3))
    assert _intervals(optimal) == _intervals(make_optimal_segments(mask, 3))
    assert len(fixed) == 3 and len(greedy) <= 3 and len(optimal) == 3


def test_optimal_segments_can_cross_zero_saving_splits():
    counts = np.array([100, 1, 1, 100])
    mask = _mask(counts.tolist())
    assert _cost(counts, make_greedy_segments(mask, 3)) == 400
    optimal = make_optimal_segments(mask, 3)
    assert len(optimal) == 3
    assert _cost(counts, optimal) == 202


def test_optimal_segments_match_bruteforce_for_small_inputs():
    rng = np.random.default_rng(0)
    for n in range(1, 9):
        for _ in range(8):
            counts = rng.integers(0, 8, size=n)
            mask = _mask(counts.tolist())
            for k in range(1, n + 1):
                seg = make_optimal_segments(mask, k)
                assert len(seg) == k
                assert _cost(counts, seg) == _brute(counts, k)


def test_solve_levels_and_schedule():
    # chain 0 -> 1 -> 2 plus independent 3
    col_rows = np.array([[1], [2], [0], [0]])
    col_mask = np.array([[True], [True], [False], [False]])
    levels = make_solve_levels(col_rows, col_mask)
    assert [lv.tolist() for lv in levels] == [[0, 3], [1], [2]]
    sched = make_level_schedule(levels, serial_width_threshold=1)
    assert sched == (("parallel", 0, 1, 0, 2), ("serial", 1, 3, 2, 0))
