"""Level sets and schedule of the triangular solves (CPU): the restatement of
core.py:488-510 / warp_backend.py:1098-1126 against brute-force properties
and hand-made known answers."""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from warpmpc_amd.problems import kkt_microbench_plan
from warpmpc_amd.qdldl import build_qdldl_plan
from warpmpc_amd.qdldl.levels import make_level_schedule, make_solve_levels, pack_levels


def test_levels_are_longest_dependency_paths():
    _, oplan = kkt_microbench_plan(horizon=10)
    q = oplan.qdldl_plan
    levels = make_solve_levels(q)
    lev = np.empty(q.n, dtype=np.int64)
    for k, cols in enumerate(levels):
        lev[cols] = k
    assert sorted(np.concatenate(levels).tolist()) == list(range(q.n))
    lp, li = np.asarray(q.lp), np.asarray(q.li)
    for c in range(q.n):
        rows = li[lp[c]:lp[c + 1]]
        assert np.all(lev[rows] > lev[c])  # a row sits above every column it depends on
    # and no lower: a row of level k > 0 has a column of level k - 1
    for r in range(q.n):
        if lev[r]:
            cols = [c for c in range(q.n) if r in li[lp[c]:lp[c + 1]]]
            assert max(lev[c] for c in cols) == lev[r] - 1
    assert len(levels) == 230 and len(levels[0]) == 184  # the mass-chain N=10 KKT


def test_diagonal_matrix_is_one_level_and_chain_is_n_levels():
    d = build_qdldl_plan(sp.csc_matrix(np.diag([1.0, 2.0, 3.0, 4.0])), ordering="natural")
    assert [c.tolist() for c in make_solve_levels(d)] == [[0, 1, 2, 3]]
    t = sp.diags([np.full(4, -1.0), np.full(5, 4.0), np.full(4, -1.0)], [-1, 0, 1], format="csc")
    chain = build_qdldl_plan(t, ordering="natural")
    assert [c.tolist() for c in make_solve_levels(chain)] == [[0], [1], [2], [3], [4]]


def test_schedule_groups_narrow_runs():
    levels = [np.arange(3), np.array([3]), np.array([4]), np.array([5, 6]), np.array([7])]
    flat, off = pack_levels(levels)
    assert flat.tolist() == list(range(8)) and off.tolist() == [0, 3, 4, 5, 7, 8]
    assert make_level_schedule(levels, off, serial_width_threshold=1) == (
        ("parallel", 0, 1, 0, 3), ("serial", 1, 3, 3, 0), ("parallel", 3, 4, 5, 2), ("serial", 4, 5, 7, 0))
    assert make_level_schedule(levels, off, serial_width_threshold=2) == (
        ("parallel", 0, 1, 0, 3), ("serial", 1, 5, 3, 0))
    assert make_level_schedule(levels, off, serial_width_threshold=3) == (("serial", 0, 5, 0, 0),)
