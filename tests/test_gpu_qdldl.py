"""GPU parity of the HIP QDLDL factor / solve against the FP64 oracle.

Error measure: per instance, max |v - v64| / |v64|_inf (normwise relative).
Tolerances (float32 kernels vs the FP64 restatement of core.py:867-1028 on the
same float32 inputs; the float32 oracle itself sits at ~5e-7 for L/D and ~7e-6
for x on these systems, whose -1/rho diagonal spans [-1000, -0.01]):
* L, D, Dinv: 1e-5;
* solutions x: 1e-4 (the bar every reported result is held to).
At full batch the same measures are taken on a random sample of instances.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from warpmpc_amd.problems import kkt_microbench_plan, kkt_microbench_values
from warpmpc_amd.qdldl import build_qdldl_plan, compile_qdldl_variant
from warpmpc_amd.qdldl import core as qcore

pytestmark = pytest.mark.gpu


LD_TOL = 1e-5
X_TOL = 1e-4


def _rel(v, ref):
    v = np.asarray(v, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.maximum(np.max(np.abs(ref), axis=1), 1e-30)
    return np.max(np.abs(v - ref), axis=1) / scale


@pytest.fixture(scope="module")
def chain10(gpu):
    _, plan = kkt_microbench_plan(horizon=10)
    return plan, compile_qdldl_variant(plan.qdldl_plan, np.float32, backend="hip", transpose_work=True)


@pytest.mark.parametrize("batch", [1, 37, 300])
def test_factor_matches_oracle(gpu, chain10, batch):
    plan, comp = chain10
    q = plan.qdldl_plan
    vals, rhs = kkt_microbench_values(plan, batch, seed=batch)
    lx, d, dinv = comp.factor(gpu.as_tensor(vals).cuda())
    lx64, d64, dinv64 = oracle.qdldl_factor(q, vals.T.astype(np.float64))
    assert _rel(lx.cpu().numpy().T, lx64).max() < LD_TOL
    assert _rel(d.cpu().numpy().T, d64).max() < LD_TOL
    assert _rel(dinv.cpu().numpy().T, dinv64).max() < LD_TOL


def test_solve_and_fused_factor_solve_match_oracle(gpu, chain10):
    plan, comp = chain10
    q = plan.qdldl_plan
    batch = 513
    vals, rhs = kkt_microbench_values(plan, batch, seed=7)
    v_t = gpu.as_tensor(vals).cuda()
    r_t = gpu.as_tensor(rhs).cuda()
    lx, d, dinv = comp.factor(v_t)
    x = comp.solve(lx, dinv, r_t).cpu().numpy().T
    x2, lx2, d2 = comp.factor_and_solve(v_t, r_t)
    x2 = x2.cpu().numpy().T
    lx64, d64, dinv64 = oracle.qdldl_factor(q, vals.T.astype(np.float64))
    x64 = oracle.qdldl_solve(q, lx64, dinv64, rhs.T.astype(np.float64))
    lx32, _, dinv32 = oracle.qdldl_factor(q, vals.T, dtype=np.float32)
    x32 = oracle.qdldl_solve(q, lx32, dinv32, rhs.T, dtype=np.float32)
    assert _rel(x, x64).max() < X_TOL
    assert _rel(x2, x64).max() < X_TOL
    assert _rel(x, x32).max() < X_TOL
    np.testing.assert_array_equal(lx2.cpu().numpy(), lx.cpu().numpy())
    np.testing.assert_array_equal(d2.cpu().numpy(), d.cpu().numpy())


def test_shared_factor_broadcasts_over_rhs(gpu, chain10):
    plan, comp = chain10
    q = plan.qdldl_plan
    vals, _ = kkt_microbench_values(plan, 1, seed=3)
    rhs = np.random.default_rng(0).standard_normal((q.n, 200)).astype(np.float32)
    lx, d, dinv = comp.factor(gpu.as_tensor(vals).cuda())
    x = comp.solve(lx, dinv, gpu.as_tensor(rhs).cuda()).cpu().numpy().T
    lx64, _, dinv64 = oracle.qdldl_factor(q, vals.T.astype(np.float64))
    x64 = oracle.qdldl_solve(q, lx64, dinv64, rhs.T.astype(np.float64))
    assert _rel(x, x64).max() < X_TOL


def test_minimal_example_on_gpu(gpu):
    import scipy.sparse as sp

    m = sp.csc_matrix([[4.0, 1.0, 0.0, 0.0], [1.0, 3.0, 0.5, 0.0], [0.0, 0.5, 2.0, 0.25],
                       [0.0, 0.0, 0.25, 1.0]])
    plan = build_qdldl_plan(m)
    # the flags of examples/jax_qdldl_minimal.py:34-42: segments accepted, validated, without
    # effect; the level-scheduled solve runs
    qcore._warned_options.clear()
    with pytest.warns(UserWarning, match="no effect"):
        comp = compile_qdldl_variant(plan, np.float32, backend="hip", transpose_work=True, segmented=True,
                                     segment_budget=4, level_scheduled_solve=True,
                                     level_scheduled_solve_threshold=2)
    assert comp.variant == "hip+transpose+levelsolve"
    assert comp.extras["options_without_effect"] == {"segmented": True}
    values = np.broadcast_to(plan.a_data.astype(np.float32), (2, plan.nnz_a)).T.copy()
    rhs = np.array([[1.0, 0.0, 0.0, 0.0], [0.0, 1.0, 2.0, 3.0]], dtype=np.float32).T.copy()
    x, _, _ = comp.factor_and_solve(gpu.as_tensor(values).cuda(), gpu.as_tensor(rhs).cuda())
    expect = np.linalg.solve(m.toarray(), rhs.astype(np.float64))
    np.testing.assert_allclose(x.cpu().numpy(), expect, rtol=2e-6, atol=1e-7)


def test_empty_batch_and_bad_shapes(gpu, chain10):
    plan, comp = chain10
    q = plan.qdldl_plan
    lx, d, dinv = comp.factor(gpu.zeros((q.nnz_a, 0), device="cuda"))
    assert lx.shape == (q.nnz_l, 0)
    with pytest.raises(ValueError):
        comp.factor(gpu.zeros((q.nnz_a + 1, 4), device="cuda"))
    with pytest.raises(ValueError):
        compile_qdldl_variant(q, np.float32, backend="warp", transpose_work=True)
    with pytest.raises(ValueError):
        compile_qdldl_variant(q, np.float32, backend="hip", transpose_work=False)


def test_microbench_pattern_full_batch_residual(gpu):
    """N = 50 chain KKT at batch 65,536: size-independent residual check on a sample."""
    _, plan = kkt_microbench_plan(horizon=50)
    q = plan.qdldl_plan
    comp = compile_qdldl_variant(q, np.float32, backend="hip", transpose_work=True)
    vals, rhs = kkt_microbench_values(plan, 65536, seed=1, device="cuda")
    x, lx, d = comp.factor_and_solve(vals, rhs)
    gpu.cuda.synchronize()
    assert bool(gpu.isfinite(x).all())
    cols = np.random.default_rng(0).choice(65536, 64, replace=False)
    v = vals[:, cols].cpu().numpy().T.astype(np.float64)
    r = rhs[:, cols].cpu().numpy().T.astype(np.float64)
    lx64, _, dinv64 = oracle.qdldl_factor(q, v)
    x64 = oracle.qdldl_solve(q, lx64, dinv64, r)
    assert _rel(x[:, cols].cpu().numpy().T, x64).max() < X_TOL
    assert _rel(lx[:, cols].cpu().numpy().T, lx64).max() < LD_TOL
