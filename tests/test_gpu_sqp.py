"""SQP step on the HIP OSQP against the float64 SQP oracle (oracle/sqp_oracle.py).

* The reference's two-stage test problem (tests/test_jax_sqp.py:62-93): the
  fixed step's shapes and finiteness (:244-270) and its values against the
  oracle; the gradient of the reference's loss through the fixed step with
  respect to the parameters (:272-323: derivatives=True, 2 refinement steps,
  beta = 0.35) against central finite differences of the float64 oracle.
* The quadrotor MPC (N = 25): one line-search step from the hover guess
  against the oracle (float32 linearization and 25 ADMM iterations against
  float64: direction within 2e-3 relative, same step length and reason),
  then 10 SQP iterations of a batch with finite iterates and a falling worst
  constraint violation.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import sqp_oracle
from parity_measure import rel_err
from warpmpc_amd.osqp import OSQPSettings
from warpmpc_amd.sqp import build_sparse_mpc_plan, compile_sparse_mpc_sqp
from warpmpc_amd.sqp.problems import quadrotor_initial_guess_and_params, quadrotor_problem, tiny_problem

pytestmark = pytest.mark.gpu


def test_tiny_fixed_step_shapes_and_values(gpu):
    prob = tiny_problem()
    s = OSQPSettings(max_iter=20, scaling=0, check_termination=0)
    plan = build_sparse_mpc_plan(prob, osqp_settings=s)
    sqp = compile_sparse_mpc_sqp(prob, plan, osqp_settings=s, segmented=True, segment_budget=2)
    z = np.array([[0.2, 0.3, 0.4, 0.5], [0.0, -0.2, 0.1, 0.0]])
    params = np.array([[1.0, 0.5], [0.2, -0.1]])
    state = sqp.init_state(2)
    result, nxt = sqp.fixed_step(z, params, beta=0.25, state=state)
    assert tuple(result.z_next.shape) == z.shape and tuple(result.solve.direction.shape) == z.shape
    assert tuple(nxt.x.shape) == z.shape and tuple(nxt.z.shape) == (2, plan.n_constraints)
    assert bool(gpu.isfinite(result.z_next).all())
    ref = sqp_oracle.OracleSQP(prob, plan).step(z, params, beta=0.25)
    assert rel_err(result.solve.direction.cpu().numpy(), ref["direction"]) < 1e-4
    assert rel_err(result.z_next.cpu().numpy(), ref["z_next"]) < 1e-4


def test_tiny_step_derivatives_match_finite_differences(gpu):
    th = gpu
    prob = tiny_problem()
    s = OSQPSettings(max_iter=80, scaling=0, check_termination=0, warm_starting=True)
    plan = build_sparse_mpc_plan(prob, osqp_settings=s, derivatives=True)
    sqp = compile_sparse_mpc_sqp(prob, plan, osqp_settings=s, derivatives=True, derivative_refinement_iters=2)
    z = np.array([[0.12, -0.08, 0.05, 0.03]])
    params0 = np.array([[0.7, -0.4]])
    weights = np.array([[1.0, -0.5, 0.25, 0.75]])
    w_t = th.as_tensor(weights, dtype=th.float32, device="cuda")
    p_t = th.as_tensor(params0, dtype=th.float32, device="cuda").requires_grad_()
    res, _ = sqp.fixed_step(z, p_t, beta=0.35, state=sqp.init_state(1))
    loss = (res.z_next * w_t).sum() + 0.1 * (res.solve.direction ** 2).sum()
    (grad,) = th.autograd.grad(loss, p_t)
    orc = sqp_oracle.OracleSQP(prob, plan)

    def loss64(p):
        out = orc.step(z, p, beta=0.35)
        return float(np.sum(out["z_next"] * weights) + 0.1 * np.sum(out["direction"] ** 2))

    eps = 1e-5
    fd = np.zeros_like(params0)
    for idx in np.ndindex(params0.shape):
        d = np.zeros_like(params0)
        d[idx] = eps
        fd[idx] = (loss64(params0 + d) - loss64(params0 - d)) / (2 * eps)
    np.testing.assert_allclose(grad.cpu().numpy(), fd, rtol=2e-3, atol=2e-5)


@pytest.fixture(scope="module")
def quad(gpu):
    prob = quadrotor_problem(horizon=25)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, scaling=0, rho_is_vec=True)
    plan = build_sparse_mpc_plan(prob, osqp_settings=s)
    return prob, plan, compile_sparse_mpc_sqp(prob, plan, osqp_settings=s)


def test_quadrotor_line_search_step_matches_oracle(gpu, quad):
    prob, plan, sqp = quad
    b = 6
    z, params = quadrotor_initial_guess_and_params(b, seed=4)
    res, _ = sqp.step(z, params, state=sqp.init_state(b))
    ref = sqp_oracle.OracleSQP(prob, plan).step(z, params)
    e = rel_err(res.solve.direction.cpu().numpy(), ref["direction"])
    print(f"quadrotor step: direction rel err {e:.2e}, steps {res.line_search.step_length.cpu().numpy()} vs "
          f"{ref['step_length']}", flush=True)
    assert e < 2e-3
    np.testing.assert_array_equal(res.line_search.step_length.cpu().numpy(), ref["step_length"])
    np.testing.assert_array_equal(res.line_search.reason.cpu().numpy(), ref["reason"])
    assert rel_err(res.z_next.cpu().numpy(), ref["z_next"]) < 2e-3


def test_quadrotor_ten_iterations(gpu, quad):
    th = gpu
    prob, plan, sqp = quad
    b = 256
    z, params = quadrotor_initial_guess_and_params(b, seed=5)
    z = th.as_tensor(z, dtype=th.float32, device="cuda")
    state = sqp.init_state(b)
    viol = []
    for it in range(10):
        res, state = sqp.step(z, params, state=state)
        z = res.z_next
        assert bool(res.is_finite.all()), it
        viol.append(float(res.line_search.constraint_violation.max()))
    print("max constraint violation per iteration:", ["%.2e" % v for v in viol], flush=True)
    _, g, lo, up = sqp.evaluate_nlp(z, params)
    scale = sqp.line_search_settings.line_search_constraint_scale
    final = scale * float(((th.clamp_min(lo - g, 0) ** 2).sum(1) + (th.clamp_max(up - g, 0) ** 2).sum(1)).sqrt().max())
    assert abs(final - viol[-1]) < 1e-6 * max(1.0, final)
    # 25 ADMM iterations per QP leave each direction inexact, so the worst
    # instance's violation falls but does not reach zero in 10 iterations
    assert final < 0.75 * viol[0]
