"""SQP line search and plan assembly (CPU only).

Pins: the known answers of the reference's line-search tests
(tests/test_jax_sqp.py:97-145: step ladder [1, 0.5, 0.25, 0.125]; the
Armijo branch accepting step 0.5 with reason COST; the baseline-cost variant
rejecting every step, reason MAX_ITER at step 0.5; violation
0.1 * sqrt(1^2 + 0.5^2)) for both the product (torch) and the oracle (numpy);
agreement of the two on random candidate sets; and the plan of the
reference's two-stage test problem (tests/test_jax_sqp.py:62-93) against its
hand-derived patterns and an independent dense assembly.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import sqp_oracle
from warpmpc_amd.sqp import (
    LINE_SEARCH_COST,
    LINE_SEARCH_MAX_ITER,
    FilterLineSearchSettings,
    build_sparse_mpc_plan,
    constraint_violation,
    filter_line_search_from_evaluations,
    make_step_lengths,
)
from warpmpc_amd.sqp.problems import quadrotor_problem, tiny_problem


def _search(settings, **kw):
    steps = make_step_lengths(settings)
    t = {k: torch.as_tensor(np.asarray(v, dtype=np.float64)) for k, v in kw.items()}
    res = filter_line_search_from_evaluations(settings=settings, step_lengths=steps, **t)
    o = sqp_oracle.filter_line_search(step_lengths=steps, baseline_cost=kw["baseline_cost"],
                                      baseline_violation=kw["baseline_constraint_violation"],
                                      armijo_metric=kw["armijo_descent_metric"],
                                      candidate_costs=kw["candidate_costs"],
                                      candidate_violations=kw["candidate_constraint_violations"],
                                      uses_trial_cost=settings.line_search_cost_accept_uses_trial_cost)
    return steps, res, o


def test_step_ladder():
    s = FilterLineSearchSettings(line_search_step_min=0.1, line_search_step_decay=0.5)
    np.testing.assert_allclose(make_step_lengths(s), [1.0, 0.5, 0.25, 0.125])
    np.testing.assert_allclose(sqp_oracle.make_step_lengths(0.5, 0.1), [1.0, 0.5, 0.25, 0.125])
    with pytest.raises(ValueError):
        make_step_lengths(FilterLineSearchSettings(line_search_step_decay=1.0))


def test_first_accepted_candidate_known_answer():
    s = FilterLineSearchSettings(line_search_step_min=0.1, line_search_step_decay=0.5)
    _, res, o = _search(s, baseline_cost=[10.0], baseline_constraint_violation=[1e-4], armijo_descent_metric=[-2.0],
                        candidate_costs=[[10.5], [9.0], [8.0]],
                        candidate_constraint_violations=[[1e-4], [1e-4], [1e-4]])
    assert float(res.step_length[0]) == 0.5 and int(res.reason[0]) == LINE_SEARCH_COST and bool(res.accepted[0])
    assert o[0][0] == 0.5 and o[1][0] == LINE_SEARCH_COST and o[2][0]


def test_baseline_cost_branch_known_answer():
    s = FilterLineSearchSettings(line_search_step_min=0.5, line_search_step_decay=0.5,
                                 line_search_cost_accept_uses_trial_cost=False)
    _, res, o = _search(s, baseline_cost=[10.0], baseline_constraint_violation=[1e-4], armijo_descent_metric=[-2.0],
                        candidate_costs=[[9.0], [8.0]], candidate_constraint_violations=[[1e-4], [1e-4]])
    assert float(res.step_length[0]) == 0.5 and int(res.reason[0]) == LINE_SEARCH_MAX_ITER
    assert not bool(res.accepted[0])
    assert o[0][0] == 0.5 and o[1][0] == LINE_SEARCH_MAX_ITER and not o[2][0]


def test_violation_known_answer():
    lower, g, upper = [[0.0, -1.0, 2.0]], [[-1.0, 0.0, 3.0]], [[2.0, 1.0, 2.5]]
    v = constraint_violation(torch.tensor(lower), torch.tensor(g), torch.tensor(upper), 0.1)
    assert abs(float(v[0]) - 0.1 * np.sqrt(1.0 + 0.25)) < 1e-7
    assert abs(sqp_oracle.constraint_violation(lower, g, upper, 0.1)[0] - 0.1 * np.sqrt(1.25)) < 1e-15


@pytest.mark.parametrize("trial", [True, False])
def test_product_and_oracle_agree_on_random_candidates(trial):
    rng = np.random.default_rng(7)
    s = FilterLineSearchSettings(line_search_cost_accept_uses_trial_cost=trial)
    k = make_step_lengths(s).size
    b = 500
    # violations on both sides of g_min / g_max, costs around the baseline
    v0 = 10.0 ** rng.uniform(-5, 0, b)
    cv = 10.0 ** rng.uniform(-5, 0, (k, b))
    c0 = rng.normal(size=b)
    cc = c0[None] + rng.normal(scale=0.01, size=(k, b))
    dm = rng.normal(size=b)
    _, res, o = _search(s, baseline_cost=c0, baseline_constraint_violation=v0, armijo_descent_metric=dm,
                        candidate_costs=cc, candidate_constraint_violations=cv)
    np.testing.assert_array_equal(res.step_length.numpy(), o[0])
    np.testing.assert_array_equal(res.reason.numpy(), o[1])
    np.testing.assert_array_equal(res.accepted.numpy(), o[2])
    # every reason occurs (with the baseline-cost Armijo test a descent step is
    # never accepted by the cost branch, so COST cannot be selected there)
    assert set(o[1].tolist()) == ({0, 1, 2, 3} if trial else {0, 1, 3})


def test_tiny_problem_plan():
    prob = tiny_problem()
    plan = build_sparse_mpc_plan(prob)
    assert (plan.n_variables, plan.n_constraints, plan.n_parameters) == (4, 3, 2)
    np.testing.assert_array_equal(plan.p_pattern.toarray() != 0, np.eye(4, dtype=bool))
    a = plan.a_pattern.toarray() != 0
    np.testing.assert_array_equal(a, [[1, 0, 1, 0], [0, 1, 0, 1], [0, 0, 1, 0]])
    np.testing.assert_array_equal(plan.representative_l, [0.0, 0.0, 0.0])
    np.testing.assert_array_equal(plan.representative_u, [1.0, 0.0, 1.0])


@pytest.mark.parametrize("which", ["tiny", "quadrotor"])
def test_sparse_assembly_matches_dense_derivatives(which):
    """The product's per-stage sparse derivatives scattered by the plan equal
    the dense derivatives of the whole NLP read at the plan's patterns."""
    prob = tiny_problem() if which == "tiny" else quadrotor_problem(horizon=3)
    plan = build_sparse_mpc_plan(prob)
    rng = np.random.default_rng(0)
    z = rng.normal(scale=0.3, size=plan.n_variables)
    p = rng.normal(scale=0.3, size=plan.n_parameters)
    (pv, av, q, l, u), _ = sqp_oracle.OracleSQP(prob, plan).linearize(z, p)
    zt, pt = torch.as_tensor(z), torch.as_tensor(p)
    pv2 = np.zeros(plan.p_pattern.nnz)
    av2 = np.zeros(plan.a_pattern.nnz)
    q2 = np.zeros(plan.n_variables)
    for st, asm in zip(prob.stages, plan.assemblies):
        zs = zt[asm.z_offset:asm.z_offset + st.z_dim][None]
        zn = zt[asm.next_z_offset:asm.next_z_offset + st.next_z_dim][None]
        ps = pt[asm.param_offset:asm.param_offset + st.param_dim][None]
        _, g, lo, up, ge, he, je = (t[0].detach().numpy() for t in st.sparse_values(zs, zn, ps))
        np.add.at(q2, asm.q_cols, ge)
        np.add.at(pv2, asm.p_pos, he)
        np.add.at(av2, asm.a_pos, je)
        rows = asm.constraint_offset + np.arange(st.constraint_dim)
        np.testing.assert_allclose(lo - g, l[rows], atol=1e-12)
        np.testing.assert_allclose(up - g, u[rows], atol=1e-12)
    np.testing.assert_allclose(q2, q, atol=1e-9)
    np.testing.assert_allclose(pv2, pv, atol=1e-9)
    np.testing.assert_allclose(av2, av, atol=1e-9)


def test_representative_point_sets_the_scaling():
    """tests/test_jax_sqp.py:226-242: with scaling on, a plan built at a
    representative point has a nontrivial D, the pattern-only plan D = 1."""
    from warpmpc_amd.osqp import OSQPSettings

    prob = tiny_problem()
    s = OSQPSettings(max_iter=10, scaling=1, check_termination=0)
    base = build_sparse_mpc_plan(prob, osqp_settings=s)
    rep = build_sparse_mpc_plan(prob, osqp_settings=s, representative_z=np.array([0.2, 0.3, 0.4, 0.5]),
                                representative_params=np.array([1.0, 0.5]))
    np.testing.assert_allclose(base.osqp_plan.scaling_d, 1.0)
    assert not np.allclose(rep.osqp_plan.scaling_d, 1.0)
    with pytest.raises(ValueError):
        build_sparse_mpc_plan(prob, osqp_settings=s, representative_z=np.zeros(4))
