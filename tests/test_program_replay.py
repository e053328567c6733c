"""CPU replay of the programs the HIP library derives (no GPU).

For each pattern a host-only plan is derived by the library itself
(wmpc_qdldl_plan_create_host), then its factor program (reach slots, column
windows, live-L slots and global L re-reads) and its chunked record sweeps are
replayed in float64 for one instance (tests/program_replay.py) and compared
with the FP64 oracle: a slot reused while live, read before written or out of
its region, or a record consuming a value outside its chunk, fails here.
Patterns cover the default LDS budget, a tiny budget that sends most live-L
re-reads to global memory, a KKT with a dense coupling row whose windows need
the narrow (32-column) LDS layout, and OSQP's adjoint matrix.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle
from program_replay import HostPlan, replay_factor_solve, replay_record_sweeps
from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan
from warpmpc_amd.problems import OSQP_INFTY, cartpole_dynamics, mass_chain_mpc, mpc_qp
from warpmpc_amd.qdldl import build_qdldl_plan


def _tiny():
    m = sp.csc_matrix([[4.0, 1.0, 0.0, 0.0], [1.0, 3.0, 0.5, 0.0], [0.0, 0.5, 2.0, 0.25],
                       [0.0, 0.0, 0.25, 1.0]])
    plan = build_qdldl_plan(m)
    return plan, plan.a_data


def _diagonal():
    plan = build_qdldl_plan(sp.diags(np.arange(1.0, 8.0), format="csc"))
    return plan, plan.a_data


def _chain_kkt():
    prob = mass_chain_mpc(horizon=10)
    op = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, OSQPSettings(rho_is_vec=True))
    return op.qdldl_plan, op.kkt_upper.data


def _dense_row_kkt():
    prob = mass_chain_mpc(horizon=10)
    a = sp.vstack([prob.a_matrix, sp.csc_matrix(np.ones((1, prob.n)))], format="csc")
    op = build_osqp_plan(prob.p_matrix, a, np.append(prob.l, -5.0), np.append(prob.u, 5.0),
                         OSQPSettings(rho_is_vec=True))
    return op.qdldl_plan, op.kkt_upper.data


def _cartpole_adjoint():
    ad, bd = cartpole_dynamics()
    prob = mpc_qp(ad, bd, 30, np.ones(4), np.ones(1), x_bound=OSQP_INFTY, u_bound=10.0)
    op = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, OSQPSettings(rho_is_vec=True),
                         derivatives=True)
    dp = op.derivative_plan
    vals = dp.adjoint_upper.data.copy()
    vals[dp.slack_diag_pos] = -np.linspace(1.0, 3.0, dp.n_ineq)   # slacks of inactive rows
    return dp.qdldl_plan, vals


CASES = {
    "tiny": (_tiny, None, 64),
    "diagonal": (_diagonal, None, 64),
    "chain_kkt": (_chain_kkt, None, 64),
    "chain_kkt_small_budget": (_chain_kkt, 24 * 1024, 64),
    "dense_row_kkt": (_dense_row_kkt, None, 32),
    "cartpole_adjoint": (_cartpole_adjoint, None, 64),
}


@pytest.mark.parametrize("name", list(CASES))
def test_programs_replay_to_the_oracle(name):
    make, budget, lcols = CASES[name]
    plan, base = make()
    hp = HostPlan(plan, lds_budget=budget)
    assert hp.slots["lds_columns"] == lcols
    rng = np.random.default_rng(len(name))
    vals = np.asarray(base, dtype=np.float64) * (1.0 + 0.05 * rng.uniform(-1, 1, size=base.shape))
    rhs = rng.standard_normal(plan.n)
    lx64, d64, dinv64 = oracle.qdldl_factor(plan, vals[None])
    x64 = oracle.qdldl_solve(plan, lx64, dinv64, rhs[None])[0]
    x, lx, d, dinv = replay_factor_solve(hp, vals, rhs)
    scale = max(1.0, np.abs(x64).max())
    np.testing.assert_allclose(lx, lx64[0], rtol=1e-9, atol=1e-9 * max(1.0, np.abs(lx64).max(initial=0)))
    np.testing.assert_allclose(d, d64[0], rtol=1e-9)
    np.testing.assert_allclose(x, x64, rtol=0, atol=1e-8 * scale)
    xs = replay_record_sweeps(hp, lx64[0], dinv64[0], rhs)
    np.testing.assert_allclose(xs, x64, rtol=0, atol=1e-8 * scale)
    if budget is not None:
        assert hp.slots["global_rereads"] > 0
