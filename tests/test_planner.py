"""Host planner: permutation, symbolic pattern and row programs (CPU only).

The planner restates warpmpc/jax_qdldl/core.py:38-311.  What pins it here:
* the L pattern equals the pattern of the exact LDLᵀ of P K Pᵀ computed densely;
* replaying the row programs (oracle factor) reconstructs P K Pᵀ = L D Lᵀ;
* solves match numpy.linalg.solve (brute force) on small systems;
* the 4x4 fixture of examples/jax_qdldl_minimal.py:22-54.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle
from warpmpc_amd.problems import kkt_microbench_plan, kkt_microbench_values, mass_chain_mpc
from warpmpc_amd.qdldl import build_qdldl_plan
from warpmpc_amd.qdldl.ordering import amd_order


def _dense_ldl(k: np.ndarray):
    n = k.shape[0]
    l = np.eye(n)
    d = np.zeros(n)
    for j in range(n):
        d[j] = k[j, j] - np.sum(l[j, :j] ** 2 * d[:j])
        for i in range(j + 1, n):
            l[i, j] = (k[i, j] - np.sum(l[i, :j] * l[j, :j] * d[:j])) / d[j]
    return l, d


def _random_quasidefinite(n1: int, n2: int, density: float, seed: int) -> sp.csc_matrix:
    rng = np.random.default_rng(seed)
    p = sp.random(n1, n1, density=density, random_state=rng)
    p = p @ p.T + sp.eye(n1)
    a = sp.random(n2, n1, density=density, random_state=rng)
    a = a + sp.eye(n2, n1)
    k = sp.bmat([[p, a.T], [a, -sp.eye(n2)]], format="csc")
    return k


def test_minimal_example_solution():
    m = sp.csc_matrix([[4.0, 1.0, 0.0, 0.0], [1.0, 3.0, 0.5, 0.0], [0.0, 0.5, 2.0, 0.25],
                       [0.0, 0.0, 0.25, 1.0]])
    plan = build_qdldl_plan(m)
    vals = np.broadcast_to(plan.a_data, (2, plan.nnz_a))
    lx, d, dinv = oracle.qdldl_factor(plan, vals)
    rhs = np.array([[1.0, 0.0, 0.0, 0.0], [0.0, 1.0, 2.0, 3.0]])
    x = oracle.qdldl_solve(plan, lx, dinv, rhs)
    np.testing.assert_allclose(x, np.linalg.solve(m.toarray(), rhs.T).T, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pattern_matches_dense_ldl_and_reconstructs(seed):
    k = _random_quasidefinite(14, 9, 0.15, seed)
    plan = build_qdldl_plan(k)
    pk = k.toarray()[np.ix_(plan.p, plan.p)]
    l_dense, d_dense = _dense_ldl(pk)
    lx, d, _ = oracle.qdldl_factor(plan, plan.a_data[None, :])
    l_sp = sp.csc_matrix((lx[0], plan.li, plan.lp), shape=(plan.n, plan.n)).toarray() + np.eye(plan.n)
    # structural pattern contains every numerical nonzero of the dense factor
    assert np.all((np.abs(l_dense) > 1e-12) <= (l_sp != 0) | np.eye(plan.n, dtype=bool))
    np.testing.assert_allclose(l_sp, l_dense, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(d[0], d_dense, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(l_sp @ np.diag(d[0]) @ l_sp.T, pk, atol=1e-10)


def test_explicit_permutation_is_used():
    k = _random_quasidefinite(8, 5, 0.3, 5)
    p = np.random.default_rng(3).permutation(13).astype(np.int32)
    plan = build_qdldl_plan(k, p=p)
    assert np.array_equal(plan.p, p)
    assert plan.ordering == "external"
    lx, d, dinv = oracle.qdldl_factor(plan, plan.a_data[None, :])
    rhs = np.arange(13, dtype=float)[None, :]
    x = oracle.qdldl_solve(plan, lx, dinv, rhs)
    np.testing.assert_allclose(x[0], np.linalg.solve(k.toarray(), rhs[0]), rtol=1e-10)
    with pytest.raises(ValueError):
        build_qdldl_plan(k, p=np.zeros(13, dtype=np.int32))


def test_amd_reduces_fill_on_mpc_kkt():
    _, plan = kkt_microbench_plan(horizon=10)
    kkt = plan.kkt_upper
    nat = build_qdldl_plan(kkt, upper=True, ordering="natural")
    amd = plan.qdldl_plan
    assert amd.nnz_l < nat.nnz_l
    # the fill of the mass-chain KKT under AMD is a small multiple of nnz(K)
    assert amd.nnz_l < 2 * plan.nnz_kkt


def test_amd_is_a_permutation_and_deterministic():
    k = _random_quasidefinite(30, 20, 0.1, 7)
    upper = sp.triu(k, format="csc")
    p1 = amd_order(50, upper.indptr, upper.indices)
    p2 = amd_order(50, upper.indptr, upper.indices)
    assert np.array_equal(p1, p2)
    assert np.array_equal(np.sort(p1), np.arange(50))


def test_mpc_kkt_solve_matches_dense_solve():
    _, plan = kkt_microbench_plan(horizon=3)
    q = plan.qdldl_plan
    vals, rhs = kkt_microbench_values(plan, 4, seed=11)
    vals = vals.T.astype(np.float64)
    rhs = rhs.T.astype(np.float64)
    lx, d, dinv = oracle.qdldl_factor(q, vals)
    x = oracle.qdldl_solve(q, lx, dinv, rhs)
    for b in range(4):
        k = sp.csc_matrix((vals[b], q.a_indices, q.a_indptr), shape=(q.n, q.n)).toarray()
        k = np.triu(k) + np.triu(k, 1).T
        np.testing.assert_allclose(x[b], np.linalg.solve(k, rhs[b]), rtol=1e-9, atol=1e-9)


def test_plan_rejects_bad_input():
    with pytest.raises(ValueError):
        build_qdldl_plan(sp.csc_matrix(np.ones((2, 3))))
    m = sp.csc_matrix(np.array([[1.0, 0.0], [0.0, 0.0]]))
    with pytest.raises(ValueError):
        build_qdldl_plan(m)  # empty column


def test_mass_chain_sizes():
    prob = mass_chain_mpc(horizon=10)
    assert prob.n == 172 and prob.m == 304
    assert prob.a_matrix.nnz == 2224
