"""Instances per wavefront (wmpc_set_instances_per_wave): every setting gives
bit-identical results.

With fewer than 64 instances per wavefront several lanes run one instance
with identical loads and stores; each instance's arithmetic is the same
instruction sequence whatever the mapping, so outputs must match bit for bit.
Batches are ragged (not a multiple of any setting) so the last wavefront
mixes real and clamped lanes.
"""

from __future__ import annotations

import numpy as np
import pytest

from warpmpc_amd import _lib
from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan, compile_osqp
from warpmpc_amd.problems import batched_mpc_data, kkt_microbench_plan, kkt_microbench_values, mass_chain_mpc
from warpmpc_amd.qdldl import compile_qdldl_variant

pytestmark = pytest.mark.gpu

SETTINGS = (8, 16, 32, 64)


@pytest.fixture
def each_ipw(gpu):
    lib = _lib.load()

    def run(fn):
        outs = {}
        try:
            for ipw in SETTINGS:
                _lib.check(lib.wmpc_set_instances_per_wave(ipw), "wmpc_set_instances_per_wave")
                outs[ipw] = [o.cpu().numpy() for o in fn()]
        finally:
            lib.wmpc_set_instances_per_wave(0)
        return outs

    return run


def _same(outs):
    base = outs[64]
    for ipw, o in outs.items():
        for a, b in zip(o, base):
            np.testing.assert_array_equal(a, b, err_msg=f"instances per wave {ipw}")


def test_auto_setting_and_bad_values(gpu):
    lib = _lib.load()
    assert lib.wmpc_instances_per_wave(65536) == 64
    assert lib.wmpc_instances_per_wave(16384) == 64
    assert lib.wmpc_instances_per_wave(4096) == 16
    assert lib.wmpc_instances_per_wave(100) == 8
    assert lib.wmpc_set_instances_per_wave(12) != 0
    assert lib.wmpc_set_instances_per_wave(0) == 0


def test_qdldl_factor_solve(each_ipw):
    _, oplan = kkt_microbench_plan(horizon=10)
    q = oplan.qdldl_plan
    comp = compile_qdldl_variant(q, np.float32, backend="hip", transpose_work=True)
    vals, rhs = kkt_microbench_values(oplan, 203, seed=3, device="cuda")

    def fn():
        lx, d, dinv = comp.factor(vals)
        x = comp.solve(lx, dinv, rhs)
        x2, lx2, d2 = comp.factor_and_solve(vals, rhs)
        return lx, dinv, x, x2, lx2, d2

    _same(each_ipw(fn))


def test_osqp_solve_shared_and_per_instance(each_ipw):
    prob = mass_chain_mpc(horizon=10)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s)
    solver = compile_osqp(plan, np.float32)
    b = 157
    pv, av, q, l, u = batched_mpc_data(prob, b, seed=4)
    rng = np.random.default_rng(4)
    pvi = pv * rng.uniform(0.5, 2.0, size=(b, pv.shape[1]))

    def fn():
        return (*solver.solve(pv, av, q, l, u)[:3], *solver.solve(pvi, av, q, l, u)[:3])

    _same(each_ipw(fn))


def test_adjoint(each_ipw):
    p = np.array([[4.0, 1.0, 0.0], [0.0, 2.0, 0.5], [0.0, 0.0, 3.0]])
    a = np.array([[1.0, 1.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    import scipy.sparse as sp

    l = np.array([1.0, -0.2, -1e30, 0.0])
    u = np.array([1.0, 0.3, 0.1, 1e30])
    plan = build_osqp_plan(sp.csc_matrix(p), sp.csc_matrix(a), l, u, OSQPSettings(max_iter=200), derivatives=True)
    solver = compile_osqp(plan, np.float32, derivatives=True, derivative_refinement_iters=20)
    b = 77
    rng = np.random.default_rng(5)
    q = rng.normal(size=(b, 3))
    x, _, y, *_ = solver.solve(plan.p_upper.data, plan.a_matrix.data, q, l, u)
    dx = rng.normal(size=(b, 3))
    dy = rng.normal(size=(b, 4))
    fused = solver.extras["adjoint_fused"]

    def fn():
        return [o for o in fused(plan.p_upper.data, plan.a_matrix.data, l, u, x, y, dx, dy, want_sol=True)]

    _same(each_ipw(fn))
