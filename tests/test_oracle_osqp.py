"""OSQP planning and the FP64 ADMM oracle (CPU only).

Pins for the oracle (restatement of warpmpc/jax_osqp/solver.py:764-872):
* the vector-rho / constraint-type known answer of tests/test_jax_osqp.py:23-50;
* the closed-form optimum x = [10, 1e-4], obj = -9.99505 of
  tests/test_jax_osqp.py:97-125 (200 iterations, Ruiz scaling 10);
* an independent dense implementation (numpy.linalg.solve on the full KKT,
  no LDLᵀ tables) of the same fixed-iteration recurrence, iterate by iterate;
* convergence to the KKT conditions of the QP on a random strictly convex QP.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from oracle import oracle
from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan
from warpmpc_amd.problems import batched_mpc_data, mass_chain_mpc


def _dense_admm(plan, p_values, a_values, q, l, u, iters, x0=None, z0=None, y0=None):
    """Same recurrence with a dense KKT solve; independent of the QDLDL tables."""
    n, m = plan.n, plan.m
    s = plan.settings
    pu = sp.csc_matrix((p_values, plan.p_upper.indices, plan.p_upper.indptr), shape=(n, n)).toarray()
    p_full = np.triu(pu) + np.triu(pu, 1).T
    a = sp.csc_matrix((a_values, plan.a_matrix.indices, plan.a_matrix.indptr), shape=(m, n)).toarray()
    d, e, c = plan.scaling_d, plan.scaling_e, plan.scaling_c
    ps = c * (d[:, None] * p_full * d[None, :])
    as_ = e[:, None] * a * d[None, :]
    kkt = np.block([[ps + s.sigma * np.eye(n), as_.T], [as_, -np.diag(plan.rho_inv_vec)]])
    qs, ls, us = q * plan.q_scale, l * e, u * e
    x = np.zeros(n) if x0 is None else x0 / d
    z = np.zeros(m) if z0 is None else z0 * e
    y = np.zeros(m) if y0 is None else y0 * plan.y_scale
    for _ in range(iters):
        rz = z - plan.rho_inv_vec * y
        sol = np.linalg.solve(kkt, np.concatenate([s.sigma * x - qs, rz]))
        zt = rz + plan.rho_inv_vec * sol[n:]
        x_new = s.alpha * sol[:n] + (1 - s.alpha) * x
        zr = s.alpha * zt + (1 - s.alpha) * z
        zn = np.minimum(np.maximum(zr + plan.rho_inv_vec * y, ls), us)
        y = y + plan.rho_vec * (zr - zn)
        x, z = x_new, zn
    x, z, y = x * d, z / e, y * plan.y_unscale
    px = p_full @ x
    prim = np.max(np.abs(a @ x - z))
    dual = np.max(np.abs(px + q + a.T @ y))
    obj = 0.5 * x @ px + q @ x
    return x, z, y, prim, dual, obj


def test_vector_rho_classification():
    p = sp.csc_matrix([[4.0, 1.0], [1.0, 2.0]])
    a = sp.csc_matrix([[1.0, 1.0], [1.0, 0.0], [0.0, 1.0], [0.0, 0.0]])
    lower = np.array([0.2, -0.1, -1e30, -1e30])
    upper = np.array([0.2, 1e30, 0.4, 1e30])
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(p, a, lower, upper, s, scaling_q=np.array([-1.0, -0.5]))
    np.testing.assert_allclose(plan.rho_vec, [100.0, 0.1, 0.1, 1e-6])
    np.testing.assert_array_equal(plan.constr_type, [1, 0, 0, -1])


def test_scaled_problem_reaches_closed_form_optimum():
    p = sp.csc_matrix([[1e-4, 0.0], [0.0, 1e4]])
    a = sp.eye(2, format="csc")
    l, u, q = np.array([0.0, 0.0]), np.array([10.0, 10.0]), np.array([-1.0, -1.0])
    plan = build_osqp_plan(p, a, l, u, OSQPSettings(max_iter=200, scaling=10), scaling_q=q)
    assert not np.allclose(plan.p_scale, 1.0) and not np.allclose(plan.a_scale, 1.0)
    x, z, y, prim, dual, obj = oracle.OracleOSQP(plan).solve(
        plan.p_upper.data[None], plan.a_matrix.data[None], q[None], l[None], u[None])
    np.testing.assert_allclose(x, [[10.0, 1e-4]], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(z, [[10.0, 1e-4]], rtol=1e-6, atol=1e-6)
    assert prim[0] < 1e-8 and dual[0] < 1e-8 and y[0, 0] > 0
    np.testing.assert_allclose(obj, [-9.99505], rtol=1e-8, atol=1e-8)


def test_oracle_matches_dense_recurrence_on_mass_chain():
    prob = mass_chain_mpc(horizon=4)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s)
    pv, av, q, l, u = batched_mpc_data(prob, 3, seed=0)
    q = q + np.random.default_rng(1).normal(size=q.shape)
    got = oracle.OracleOSQP(plan).solve(pv, av, q, l, u)
    for b in range(3):
        ref = _dense_admm(plan, pv[0], av[0], q[b], l[b], u[b], 25)
        for g, r in zip(got, ref):
            np.testing.assert_allclose(np.asarray(g)[b], r, rtol=1e-8, atol=1e-9)


def test_oracle_matches_dense_recurrence_with_scaling_and_warm_start():
    rng = np.random.default_rng(4)
    n, m = 6, 9
    m0 = rng.normal(size=(n, n))
    p = sp.csc_matrix(m0 @ m0.T + np.eye(n))
    a = sp.csc_matrix(np.vstack([rng.normal(size=(3, n)), np.eye(n)]))
    l = np.concatenate([-np.ones(3), -0.5 * np.ones(n)])
    u = np.concatenate([np.ones(3), 0.5 * np.ones(n)])
    l[0] = u[0] = 0.3
    q = rng.normal(size=n)
    s = OSQPSettings(max_iter=15, scaling=5, rho_is_vec=True)
    plan = build_osqp_plan(p, a, l, u, s, scaling_q=q)
    x0, z0, y0 = rng.normal(size=n), rng.normal(size=m), rng.normal(size=m)
    got = oracle.OracleOSQP(plan).solve(plan.p_upper.data[None], plan.a_matrix.data[None], q[None],
                                        l[None], u[None], x0=x0[None], z0=z0[None], y0=y0[None])
    ref = _dense_admm(plan, plan.p_upper.data, plan.a_matrix.data, q, l, u, 15, x0, z0, y0)
    for g, r in zip(got, ref):
        np.testing.assert_allclose(np.asarray(g)[0], r, rtol=1e-8, atol=1e-9)


def test_converged_oracle_satisfies_qp_optimality():
    rng = np.random.default_rng(9)
    n = 5
    m0 = rng.normal(size=(n, n))
    p = sp.csc_matrix(m0 @ m0.T + 0.5 * np.eye(n))
    a = sp.csc_matrix(np.vstack([np.ones((1, n)), np.eye(n)]))
    l = np.concatenate([[1.0], -0.3 * np.ones(n)])
    u = np.concatenate([[1.0], 0.3 * np.ones(n)])
    q = rng.normal(size=n)
    plan = build_osqp_plan(p, a, l, u, OSQPSettings(max_iter=4000, rho_is_vec=True))
    x, z, y, prim, dual, obj = oracle.OracleOSQP(plan).solve(
        plan.p_upper.data[None], plan.a_matrix.data[None], q[None], l[None], u[None])
    x, y = x[0], y[0]
    ax = a.toarray() @ x
    assert prim[0] < 1e-9 and dual[0] < 1e-9
    assert np.all(ax >= l - 1e-9) and np.all(ax <= u + 1e-9)
    # complementarity: y_i > 0 only at the upper bound, < 0 only at the lower bound
    assert np.all((y <= 1e-8) | (np.abs(ax - u) < 1e-7))
    assert np.all((y >= -1e-8) | (np.abs(ax - l) < 1e-7))
    np.testing.assert_allclose(p.toarray() @ x + q + a.toarray().T @ y, 0.0, atol=1e-8)


def test_nan_input_propagates_and_stays_per_instance():
    """A NaN in one instance's q reaches its outputs (the box projection and the
    residual maxima propagate NaN instead of clamping it away) and leaves the
    other instances untouched."""
    p = sp.csc_matrix([[4.0, 1.0], [1.0, 2.0]])
    a = sp.csc_matrix([[1.0, 1.0], [1.0, 0.0], [0.0, 1.0]])
    l = np.array([1.0, 0.0, 0.0])
    u = np.array([1.0, 0.7, 0.7])
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=50)
    plan = build_osqp_plan(sp.triu(p, format="csc"), a, l, u, s)
    orc = oracle.OracleOSQP(plan)
    q = np.array([[1.0, 1.0], [np.nan, 1.0], [1.0, 1.0]])
    x, z, y, prim, dual, obj = orc.solve(plan.p_upper.data[None], plan.a_matrix.data[None], q, l, u)
    assert np.isnan(x[1]).all() and np.isnan(z[1]).all() and np.isnan(y[1]).all()
    assert np.isnan(prim[1]) and np.isnan(dual[1]) and np.isnan(obj[1])
    for i in (0, 2):
        assert np.isfinite(x[i]).all() and np.isfinite(prim[i])
    np.testing.assert_array_equal(x[0], x[2])
