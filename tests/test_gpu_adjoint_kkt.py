"""GPU parity of the KKT form of the OSQP adjoint (derivative_mode="kkt").

The KKT form differentiates the active-set KKT system; where the rows with a
positive multiplier are feasible (a converged solution) that is the limit the
reference's adjoint refinement (warpmpc/jax_osqp/solver.py:1047-1139)
converges to.  Its oracle is ``oracle.adjoint_exact(..., clamp_slacks=True)``
(the reference's formula with M0 sol = rhs solved exactly, densely in
float64, slacks of rows with a positive multiplier taken as min(s, 0)),
pinned on the CPU against the C restatement run to convergence, the dense
active-set implicit derivative (also at infeasible iterates) and finite
differences (tests/test_oracle_adjoint.py).

* Pinned small QPs (equality rows of both multiplier signs, upper-, lower-
  and two-sided rows), converged FP64 (x, y) rounded to float32: every output
  (sol, dq, dl, du, dP, dA) within 1e-4 of the oracle, error per instance
  norm-wise relative (parity_measure.adjoint_errors).
* The cart-pole workload (N=30, batch 16384, x and y from the 25-iteration
  HIP forward): within 1e-4 of the oracle on sampled instances, where the
  reference's own algorithm at its default 100 refinement steps lies several
  percent from the same limit.
* End to end: d loss / d(theta, x0) against FP64 central finite differences.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle
from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan, compile_osqp
from test_gpu_adjoint import (HORIZON, SMALL, TIGHT, _bounds, _cartpole_batch, _errs, _fused, _p_values,
                              _small)

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _random_qp(seed):
    rng = np.random.default_rng(seed)
    n, m = 6, 9
    g = rng.standard_normal((n, n))
    p = g @ g.T + n * np.eye(n)
    p[np.abs(p) < 0.8] = 0.0
    p = p + np.eye(n)
    a = rng.standard_normal((m, n))
    a[rng.random((m, n)) < 0.4] = 0.0
    a[np.arange(m), rng.integers(0, n, m)] = 1.0
    q = rng.standard_normal(n) * 3
    l = -rng.random(m) * 0.5
    u = rng.random(m) * 0.5
    l[0] = u[0] = 0.2
    l[1] = -1e30
    u[2] = 1e30
    return sp.csc_matrix(np.triu(p)), sp.csc_matrix(a), l, u, q


@pytest.mark.parametrize("name", list(SMALL) + ["random_3", "random_11"])
def test_kkt_form_matches_exact_oracle_on_pinned_cases(gpu, name):
    p, a, l, u, q = _random_qp(int(name.split("_")[1])) if name.startswith("random") else _small(name)
    plan = build_osqp_plan(p, a, l, u, TIGHT, derivatives=True)
    pv, av = plan.p_upper.data[None], plan.a_matrix.data[None]
    x, _, y, *_ = oracle.OracleOSQP(plan).solve(pv, av, q[None], l[None], u[None])
    rng = np.random.default_rng(4)
    b = 70
    xs = np.repeat(x, b, 0).astype(np.float32).astype(np.float64)
    ys = np.repeat(y, b, 0).astype(np.float32).astype(np.float64)
    dx = rng.standard_normal((b, plan.n)).astype(np.float32).astype(np.float64)
    dy = rng.standard_normal((b, plan.m)).astype(np.float32).astype(np.float64)
    print(f"[kkt adjoint] {name}: n={plan.n} m={plan.m} batch {b}", flush=True)
    solver = compile_osqp(plan, np.float32, derivatives=True, derivative_mode="kkt")
    assert solver.extras["adjoint_plan"].form == "kkt"
    got = _fused(solver, pv, av, l, u, xs, ys, dx, dy, want_sol=True)
    ref = oracle.adjoint_exact(plan, pv, av, l, u, xs, ys, dx, dy, clamp_slacks=True)
    errs = _errs(got, (*ref[3:], ref[0]), ("dq", "dl", "du", "dp", "da", "sol"))
    assert max(errs.values()) < TOL, errs


@pytest.mark.parametrize("seed", [3, 11, 19])
def test_kkt_form_on_an_infeasible_iterate(gpu, seed):
    """(x, y) after 3 ADMM iterations: rows with a positive multiplier violate
    their bound (slack s > 0); the KKT form takes their slack as 0 and matches
    the active-set oracle."""
    p, a, l, u, q = _random_qp(seed)
    plan = build_osqp_plan(p, a, l, u, OSQPSettings(max_iter=3, rho=0.1, sigma=1e-6, alpha=1.6), derivatives=True)
    pv, av = plan.p_upper.data[None], plan.a_matrix.data[None]
    x, _, y, *_ = oracle.OracleOSQP(plan).solve(pv, av, q[None], l[None], u[None])
    ax = a.toarray() @ x[0]
    two = l < u
    bad = two & (((y[0] > 0) & (ax > u)) | ((y[0] < 0) & (ax < l)))
    assert bad.any()
    rng = np.random.default_rng(seed)
    b = 64
    xs = np.repeat(x, b, 0).astype(np.float32).astype(np.float64)
    ys = np.repeat(y, b, 0).astype(np.float32).astype(np.float64)
    dx = rng.standard_normal((b, plan.n)).astype(np.float32).astype(np.float64)
    dy = rng.standard_normal((b, plan.m)).astype(np.float32).astype(np.float64)
    print(f"[kkt adjoint] 3-iteration iterate of random_{seed}: {int(bad.sum())} violated rows with a multiplier",
          flush=True)
    solver = compile_osqp(plan, np.float32, derivatives=True, derivative_mode="kkt")
    got = _fused(solver, pv, av, l, u, xs, ys, dx, dy, want_sol=True)
    ref = oracle.adjoint_exact(plan, pv, av, l, u, xs, ys, dx, dy, clamp_slacks=True)
    errs = _errs(got[:5], ref[3:])
    assert max(errs.values()) < TOL, errs


@pytest.fixture(scope="module")
def cartpole_kkt(gpu):
    prob, *_ = _cartpole_batch(1, 0)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s, derivatives=True)
    return prob, plan, compile_osqp(plan, np.float32, derivatives=True, derivative_mode="kkt")


@pytest.mark.parametrize("u_bound", [10.0, 0.3])
def test_kkt_form_cartpole_workload_matches_exact_oracle(gpu, u_bound):
    th = gpu
    b = 16384
    prob, tq, tr, x0 = _cartpole_batch(b, 1, u_bound=u_bound)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s, derivatives=True)
    solver = compile_osqp(plan, np.float32, derivatives=True, derivative_mode="kkt")
    pv = _p_values(tq, tr)
    av = plan.a_matrix.data[None]
    l, u = _bounds(prob, x0)
    print(f"[kkt adjoint] cart-pole |u| <= {u_bound}, batch {b}", flush=True)
    x, _, y, *_ = solver.solve(pv, av, np.zeros((b, plan.n)), l, u)
    rng = np.random.default_rng(101)
    dx = th.as_tensor(rng.standard_normal((b, plan.n)), dtype=th.float32, device="cuda")
    dy = th.as_tensor(rng.standard_normal((b, plan.m)), dtype=th.float32, device="cuda")
    got = _fused(solver, pv, av, l, u, x, y, dx, dy)[:5]
    assert all(np.isfinite(g).all() for g in got)
    idx = np.unique(np.concatenate([rng.choice(b, 64, replace=False), [0, 63, 64, b - 1]]))
    f64 = lambda t: t.cpu().numpy().astype(np.float64)[idx]  # noqa: E731
    args = (pv.astype(np.float32).astype(np.float64)[idx], av, l.astype(np.float32).astype(np.float64)[idx],
            u.astype(np.float32).astype(np.float64)[idx], f64(x), f64(y), f64(dx), f64(dy))
    ref = oracle.adjoint_exact(plan, *args, clamp_slacks=True)[3:]
    r101 = oracle.OracleAdjoint(plan, refine=100).compute(*args)[3:]
    active = (np.abs(f64(y)) > 0).sum(1)
    print(f"  multipliers nonzero per instance: {active.min()}..{active.max()} (124 equality rows)", flush=True)
    errs, errs101 = _errs([g[idx] for g in got], ref), _errs(r101, ref)
    for name_ in errs:
        e, e101 = errs[name_], errs101[name_]
        print(f"  {name_}: KKT form vs exact {e:.2e}   reference algorithm (R=101, FP64) vs the same limit {e101:.2e}",
              flush=True)
        assert e < TOL, (name_, e)


def test_kkt_form_gradient_matches_finite_differences(gpu):
    """d/d(theta, x0) of sum_b x0ᵀ Q x0 + R u0² (|u| <= 2, some inputs on
    their bounds) with the forward run to convergence."""
    th = gpu
    b = 4
    prob, tq, tr, x0 = _cartpole_batch(b, 5, u_bound=2.0)
    conv = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=4000, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, conv, derivatives=True)
    solver = compile_osqp(plan, np.float32, derivatives=True, derivative_mode="kkt")
    ui = (HORIZON + 1) * 4

    def loss_fd(tq_, tr_, x0_):
        pv = _p_values(tq_, tr_)
        l, u = _bounds(prob, x0_)
        x, *_ = oracle.OracleOSQP(plan, max_iter=10000).solve(pv, plan.a_matrix.data[None],
                                                               np.zeros((b, plan.n)), l, u)
        return np.sum(np.exp(tq_) * x0_ ** 2, 1) + np.exp(tr_[:, 0]) * x[:, ui] ** 2

    tq_t = th.as_tensor(tq, dtype=th.float32, device="cuda").requires_grad_()
    tr_t = th.as_tensor(tr, dtype=th.float32, device="cuda").requires_grad_()
    x0_t = th.as_tensor(x0, dtype=th.float32, device="cuda").requires_grad_()
    pv = _p_values(tq_t, tr_t, lib=th)
    l0, u0 = (th.as_tensor(v, dtype=th.float32, device="cuda") for v in _bounds(prob, x0))
    l = th.cat([-x0_t, l0[:, 4:]], 1)
    u = th.cat([-x0_t, u0[:, 4:]], 1)
    x = solver.solve_primal(pv, plan.a_matrix.data[None], th.zeros((1, plan.n), device="cuda"), l, u)
    loss = (th.exp(tq_t) * x0_t ** 2).sum() + (th.exp(tr_t[:, 0]) * x[:, ui] ** 2).sum()
    grads = [g.cpu().numpy().astype(np.float64) for g in th.autograd.grad(loss, (tq_t, tr_t, x0_t))]
    h = 1e-6
    base = [tq, tr, x0]
    for which, got in enumerate(grads):
        arr = base[which]
        fd = np.zeros_like(arr)
        for k in range(arr.shape[1]):
            d = np.zeros_like(arr)
            d[:, k] = h
            plus, minus = list(base), list(base)
            plus[which] = arr + d
            minus[which] = arr - d
            fd[:, k] = (loss_fd(*plus) - loss_fd(*minus)) / (2 * h)
        scale = np.maximum(np.abs(fd).max(), 1e-3)
        assert np.abs(got - fd).max() / scale < 2e-3, (which, got, fd)


def test_kkt_form_state_api_and_shared_operands(gpu, cartpole_kkt):
    prob, plan, solver = cartpole_kkt
    th = gpu
    b = 100
    _, tq, tr, x0 = _cartpole_batch(b, 2)
    pv = _p_values(tq, tr)
    l, u = _bounds(prob, x0)
    x, _, y, *_ = solver.solve(pv, plan.a_matrix.data[None], np.zeros((b, plan.n)), l, u)
    dx = th.randn((b, plan.n), device="cuda")
    dy = th.randn((b, plan.m), device="cuda")
    state = solver.adjoint_derivative_compute(pv, plan.a_matrix.data[None], l, u, x, y, dx, dy)
    dq, dl, du = solver.adjoint_derivative_get_vec(state)
    fq, fl, fu, fp, fa, fsol = _fused(solver, pv, plan.a_matrix.data[None], l, u, x, y, dx, dy, want_sol=True)
    np.testing.assert_array_equal(state.sol.cpu().numpy(), fsol)
    for got, ref in ((dq, fq), (dl, fl), (du, fu)):
        np.testing.assert_array_equal(got.cpu().numpy(), ref)
    # A shared by the batch (stride 0) and A repeated per instance agree exactly
    per = _fused(solver, pv, np.repeat(plan.a_matrix.data[None], b, 0), l, u, x, y, dx, dy)
    for s_, p_ in zip((fq, fl, fu, fp, fa), per):
        np.testing.assert_array_equal(s_, p_)


def test_derivative_mode_is_validated(gpu, cartpole_kkt):
    prob, plan, _ = cartpole_kkt
    with pytest.raises(ValueError, match="derivative_mode"):
        compile_osqp(plan, np.float32, derivatives=True, derivative_mode="dense")


def test_kkt_form_on_the_quadrotor_sqp_qp(gpu):
    """The QP of the nonlinear quadrotor MPC after three SQP iterations (312
    equality rows, thrust bounds active): du0/dx0 through the KKT form within
    1e-4 of the FP64 oracle.  Without the per-instance equilibration, the
    float32 factorization's element growth on this equality-heavy KKT makes
    the refinement diverge."""
    th = gpu
    from warpmpc_amd.sqp import build_sparse_mpc_plan, compile_sparse_mpc_sqp
    from warpmpc_amd.sqp.problems import QUAD_NX, quadrotor_initial_guess_and_params, quadrotor_problem

    prob = quadrotor_problem(horizon=25)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, scaling=0, rho_is_vec=True)
    plan = build_sparse_mpc_plan(prob, osqp_settings=s, derivatives=True)
    sqp = compile_sparse_mpc_sqp(prob, plan, osqp_settings=s, derivatives=True, derivative_mode="kkt")
    op = plan.osqp_plan
    b = 64
    z, params = quadrotor_initial_guess_and_params(b, seed=8)
    z = th.as_tensor(z, dtype=th.float32, device="cuda")
    state = sqp.init_state(b)
    for _ in range(3):
        res, state = sqp.step(z, params, state=state)
        z = res.z_next
    lin = sqp.linearize_nlp(z, params)
    d, _, y, *_ = sqp.osqp.solve_warm_start(lin.p_values, lin.a_values, lin.q, lin.l, lin.u, *state)
    fused = sqp.osqp.extras["adjoint_fused"]
    f64 = lambda t: t.cpu().numpy().astype(np.float64)  # noqa: E731
    worst = 0.0
    print(f"[kkt adjoint] quadrotor SQP QP (n={op.n}, m={op.m}), batch {b}", flush=True)
    for k in range(4):
        dx = th.zeros_like(d)
        dx[:, QUAD_NX + k] = 1.0
        got = fused(lin.p_values, lin.a_values, lin.l, lin.u, d, y, dx, None, want_mat=True)
        ref = oracle.adjoint_exact(op, f64(lin.p_values), f64(lin.a_values), f64(lin.l), f64(lin.u), f64(d), f64(y),
                                   f64(dx), np.zeros((b, op.m)), clamp_slacks=True)
        errs = _errs([f64(g) for g in got[:5]], ref[3:])
        worst = max(worst, *errs.values())
        assert max(errs.values()) < TOL, (k, errs)
    print(f"  worst output error {worst:.2e}", flush=True)
