"""Segment planner (the ``segmented`` option), CPU only.

Pins: the known answers of the reference's segment tests
(tests/test_jax_qdldl_segments.py:56-88: row widths [100, 1, 1, 100] in three
segments cost 400 padded slots greedily and 202 optimally), exhaustive search
over every split for n <= 8, and the partition invariants every strategy must
keep.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest

from warpmpc_amd.qdldl.segments import make_segments

STRATEGIES = ("fixed", "greedy", "optimal")


def rows_of_width(widths):
    """Row mask whose row i holds widths[i] entries."""
    widths = list(widths)
    mask = np.zeros((len(widths), max(widths, default=0)), dtype=bool)
    for i, w in enumerate(widths):
        mask[i, :w] = True
    return mask


def bounds(segments):
    return [(int(cols[0]), int(cols[-1]) + 1) for _, cols in segments]


def padded_cost(widths, segments):
    widths = np.asarray(widths)
    return sum((b - a) * int(widths[a:b].max(initial=0)) for a, b in bounds(segments))


def cheapest_split(widths, pieces):
    n = len(widths)
    costs = []
    for cuts in itertools.combinations(range(1, n), pieces - 1):
        edges = (0, *cuts, n)
        costs.append(sum((b - a) * max(widths[a:b], default=0) for a, b in zip(edges, edges[1:])))
    return min(costs)


@pytest.mark.parametrize("strategy, expected", [("greedy", 400), ("optimal", 202)])
def test_reference_known_answer(strategy, expected):
    widths = [100, 1, 1, 100]
    segs = make_segments(rows_of_width(widths), segment_budget=3, segment_strategy=strategy)
    assert padded_cost(widths, segs) == expected


@pytest.mark.parametrize("n", range(1, 9))
def test_optimal_equals_exhaustive_search(n):
    rng = np.random.default_rng(100 + n)
    for _ in range(6):
        widths = rng.integers(0, 9, size=n).tolist()
        for pieces in range(1, n + 1):
            segs = make_segments(rows_of_width(widths), segment_budget=pieces, segment_strategy="optimal")
            assert len(segs) == pieces
            assert padded_cost(widths, segs) == cheapest_split(widths, pieces)


@pytest.mark.parametrize("strategy", STRATEGIES)
@pytest.mark.parametrize("budget", [1, 2, 5, 64])
def test_segments_partition_the_rows(strategy, budget):
    widths = np.random.default_rng(budget).integers(0, 20, size=37).tolist()
    segs = make_segments(rows_of_width(widths), segment_budget=budget, segment_strategy=strategy)
    edges = bounds(segs)
    assert edges[0][0] == 0 and edges[-1][1] == len(widths)
    assert all(a < b for a, b in edges)
    assert all(edges[i][1] == edges[i + 1][0] for i in range(len(edges) - 1))
    assert len(segs) <= min(budget, len(widths))
    for start, cols in segs:
        assert start == cols[0] and np.array_equal(cols, np.arange(cols[0], cols[-1] + 1))
    # no strategy can do better than the exact minimum, nor worse than one segment
    best = padded_cost(widths, make_segments(rows_of_width(widths), segment_budget=budget,
                                             segment_strategy="optimal"))
    assert best <= padded_cost(widths, segs) <= len(widths) * max(widths)


def test_empty_pattern_and_bad_options():
    assert make_segments(rows_of_width([]), segment_budget=4, segment_strategy="optimal") == []
    with pytest.raises(ValueError):
        make_segments(rows_of_width([1, 2]), segment_budget=0, segment_strategy="greedy")
    with pytest.raises(ValueError):
        make_segments(rows_of_width([1, 2]), segment_budget=2, segment_strategy="balanced")
