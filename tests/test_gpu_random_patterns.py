"""GPU parity on random sparsity patterns (factor, solve, fused OSQP).

The MPC patterns of the other tests are banded; these are random quasi-
definite KKT matrices and random sparse QPs from 3 to 160 rows, so that the
programs' tables range from a few entries to many 64-entry chunks and the
reach / update runs take every length.  Each case is measured against the FP64
oracle and held to the larger of the stated tolerance and four times the error
of the float32 restatement on the same inputs (a random instance may be worse
conditioned than the MPC ones).
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle
from parity_measure import osqp_errors, rel_err
from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan, compile_osqp
from warpmpc_amd.qdldl import build_qdldl_plan, compile_qdldl_variant

pytestmark = pytest.mark.gpu

CASES = [(3, 2, 0.5, 0), (17, 23, 0.2, 1), (40, 60, 0.1, 2), (90, 70, 0.05, 3), (30, 5, 0.6, 4)]
BATCH = 130  # two wavefronts and a ragged third


def _random_pd(rng, n, density):
    g = sp.random(n, n, density=density, random_state=rng, data_rvs=rng.standard_normal)
    return (g @ g.T + sp.eye(n) * (1.0 + n * density)).tocsc()


def _kkt_case(n, m, density, seed):
    """Upper triangle of [[P, Aᵀ], [A, -diag(r)]] and per-instance values that
    keep it quasi-definite (P scaled by a positive factor with a diagonal
    jitter, A entries scaled, r in [0.1, 10])."""
    rng = np.random.default_rng(seed)
    p = _random_pd(rng, n, density)
    a = sp.random(m, n, density=density, random_state=rng, data_rvs=rng.standard_normal).tocsc()
    a = (a + sp.csc_matrix((np.ones(m), (np.arange(m), rng.integers(0, n, m))), shape=(m, n))).tocsc()
    k = sp.bmat([[p, a.T], [a, -sp.eye(m)]], format="csc")
    ku = sp.triu(k, format="csc")
    ku.sort_indices()
    rows = ku.indices
    cols = np.repeat(np.arange(n + m), np.diff(ku.indptr))
    in_p = (rows < n) & (cols < n)
    diag_p = in_p & (rows == cols)
    diag_c = (rows >= n) & (rows == cols)
    vals = np.repeat(ku.data[None, :], BATCH, 0)
    vals[:, in_p] *= rng.uniform(0.5, 2.0, size=(BATCH, 1))
    vals[:, diag_p] += rng.uniform(0.0, 1.0, size=(BATCH, int(diag_p.sum())))
    off = ~in_p & ~diag_c
    vals[:, off] *= rng.uniform(0.5, 1.5, size=(BATCH, int(off.sum())))
    vals[:, diag_c] = -rng.uniform(0.1, 10.0, size=(BATCH, int(diag_c.sum())))
    rhs = rng.standard_normal((BATCH, n + m))
    return ku, vals.astype(np.float32).astype(np.float64), rhs.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("n,m,density,seed", CASES)
def test_random_kkt_factor_and_solve(gpu, n, m, density, seed):
    th = gpu
    ku, vals, rhs = _kkt_case(n, m, density, seed)
    plan = build_qdldl_plan(ku, upper=True)
    comp = compile_qdldl_variant(plan, np.float32, backend="hip", transpose_work=True)
    print(f"[random kkt] n_K={n + m} nnz={ku.nnz} nnz_L={plan.nnz_l} seed {seed}", flush=True)
    f32 = dict(dtype=th.float32, device="cuda")
    v_t = th.as_tensor(vals.T.copy(), **f32)
    r_t = th.as_tensor(rhs.T.copy(), **f32)
    x, lx, d = comp.factor_and_solve(v_t, r_t)
    lx2, d2, dinv2 = comp.factor(v_t)
    x2 = comp.solve(lx2, dinv2, r_t)
    lx64, d64, dinv64 = oracle.qdldl_factor(plan, vals)
    x64 = oracle.qdldl_solve(plan, lx64, dinv64, rhs)
    lx32, d32, dinv32 = oracle.qdldl_factor(plan, vals, dtype=np.float32)
    x32 = oracle.qdldl_solve(plan, lx32, dinv32, rhs.astype(np.float32), dtype=np.float32)
    for name, got, ref, ref32, tol in (("L", lx, lx64, lx32, 1e-5), ("D", d, d64, d32, 1e-5),
                                       ("x", x, x64, x32, 1e-4), ("L (factor)", lx2, lx64, lx32, 1e-5),
                                       ("x (solve)", x2, x64, x32, 1e-4)):
        e = rel_err(got.cpu().numpy().T, ref)
        e32 = rel_err(ref32, ref)
        print(f"  {name}: {e:.2e} (float32 oracle {e32:.2e})", flush=True)
        assert e < max(tol, 4 * e32), (name, e, e32)
    # the fused and the separate kernels run the same operations
    np.testing.assert_array_equal(lx.cpu().numpy(), lx2.cpu().numpy())


@pytest.mark.parametrize("n,m,density,seed", CASES)
@pytest.mark.parametrize("per_instance", [False, True])
def test_random_qp_osqp(gpu, n, m, density, seed, per_instance):
    rng = np.random.default_rng(100 + seed)
    p = sp.triu(_random_pd(rng, n, density), format="csc")
    a = sp.random(m, n, density=density, random_state=rng, data_rvs=rng.standard_normal).tocsc()
    a = (a + sp.csc_matrix((np.ones(m), (np.arange(m), rng.integers(0, n, m))), shape=(m, n))).tocsc()
    l = -rng.uniform(0.1, 1.0, m)
    u = rng.uniform(0.1, 1.0, m)
    eq = rng.random(m) < 0.2
    l[eq] = u[eq] = rng.uniform(-0.5, 0.5, int(eq.sum()))
    u[(~eq) & (rng.random(m) < 0.2)] = 1e30
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(p, a, l, u, s)
    solver = compile_osqp(plan, np.float32, specialize=False)
    pv, av = plan.p_upper.data[None], plan.a_matrix.data[None]
    if per_instance:
        pv = pv * rng.uniform(0.5, 2.0, size=(BATCH, 1))
        av = av * rng.uniform(0.5, 1.5, size=(BATCH, av.shape[1]))
    q = rng.standard_normal((BATCH, n))
    lb = np.repeat(l[None], BATCH, 0)
    ub = np.repeat(u[None], BATCH, 0)
    pv, av, q = (v.astype(np.float32).astype(np.float64) for v in (pv, av, q))
    print(f"[random qp] n={n} m={m} nnz_L={plan.qdldl_plan.nnz_l} per_instance={per_instance}", flush=True)
    out = [o.cpu().numpy() for o in solver.solve(pv, av, q, lb, ub)]
    ref = oracle.OracleOSQP(plan).solve(pv, av, q, lb, ub)
    ref32 = oracle.OracleOSQP(plan).solve(pv, av, q, lb, ub, dtype=np.float32)
    errs = osqp_errors(out, ref, q)
    errs32 = osqp_errors(ref32, ref, q)
    print("  " + " ".join(f"{k} {v:.1e} (fp32 {errs32[k]:.1e})" for k, v in errs.items()), flush=True)
    bad = {k: v for k, v in errs.items() if not v < max(1e-4, 4 * errs32[k])}
    assert not bad, (bad, errs32)
