"""Heterogeneous batch in one persistent launch (wmpc_osqp_solve_hetero_f32).

The nine patterns of the heterogeneous workload ((nx, nu) in {(4,2), (12,4),
(24,8)} x N in {10, 20, 40}) solved together:
* at small ragged batches, every pattern against the FP64 ADMM oracle
  (parity_measure.osqp_errors < 1e-4: norm-wise relative per instance, 25
  float32 iterations) and bit for bit against that pattern's own single-pattern
  launch with per-instance factors (same per-instance arithmetic);
* warm-started from the previous solution shifted one stage, against the
  oracle given the same warm start;
* the full 200,000-instance batch (instances drawn uniformly over the
  patterns): every output finite and sampled instances against the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from parity_measure import osqp_errors, rel_err
from warpmpc_amd.osqp import HeterogeneousOSQP, OSQPSettings, ShiftMaps, build_osqp_plan, compile_osqp, \
    shift_warm_start
from warpmpc_amd.problems import batched_mpc_data, hetero_counts, hetero_mpc_patterns, mpc_shift_maps

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def nine(gpu):
    pats = hetero_mpc_patterns()
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plans = [build_osqp_plan(p.p_matrix, p.a_matrix, p.l, p.u, s) for p in pats]
    solvers = [compile_osqp(pl, np.float32) for pl in plans]
    return pats, plans, solvers, HeterogeneousOSQP(solvers)


def _data(pats, counts, seed):
    return [batched_mpc_data(p, int(c), seed=seed + k) for k, (p, c) in enumerate(zip(pats, counts))]


def test_every_pattern_matches_oracle_and_single_pattern_launch(gpu, nine):
    pats, plans, solvers, het = nine
    counts = [37, 64, 70, 129, 5, 1, 33, 90, 17]
    data = _data(pats, counts, 10)
    print("hetero: batches", counts, flush=True)
    outs = het.solve(data)
    for k, (plan, solver, d, out) in enumerate(zip(plans, solvers, data, outs)):
        ref = oracle.OracleOSQP(plan).solve(*d)
        errs = osqp_errors([o.cpu().numpy() for o in out], ref, d[2])
        print(f"  pattern {k}:", " ".join(f"{n} {e:.1e}" for n, e in errs.items()), flush=True)
        assert max(errs.values()) < TOL, (k, errs)
        b = counts[k]
        pv = np.broadcast_to(d[0], (b, d[0].shape[1]))
        av = np.broadcast_to(d[1], (b, d[1].shape[1]))
        single = solver.solve(pv, av, *d[2:])  # per-instance matrices: per-instance factor
        for o, s_ in zip(out, single):
            np.testing.assert_array_equal(o.cpu().numpy(), s_.cpu().numpy())


def test_shifted_warm_start_matches_oracle(gpu, nine):
    pats, plans, solvers, het = nine
    th = gpu
    counts = [20, 3, 65, 8, 40, 12, 9, 30, 2]
    data = _data(pats, counts, 20)
    first = het.solve(data)
    warm = []
    for p, plan, o in zip(pats, plans, first):
        var, con = mpc_shift_maps(p)
        maps = ShiftMaps(var, con, plan.n, plan.m)
        x, z, y = (t.t().contiguous() for t in o[:3])
        sx, sz, sy = shift_warm_start(maps, x, z, y)
        np.testing.assert_array_equal(sx.cpu().numpy(), x.cpu().numpy()[var])
        np.testing.assert_array_equal(sy.cpu().numpy(), y.cpu().numpy()[con])
        warm.append(tuple(t.t().contiguous() for t in (sx, sz, sy)))
    outs = het.solve(data, warm_starts=warm)
    for k, (plan, d, w, out) in enumerate(zip(plans, data, warm, outs)):
        ws = [t.cpu().numpy().astype(np.float64) for t in w]
        ref = oracle.OracleOSQP(plan).solve(*d, x0=ws[0], z0=ws[1], y0=ws[2])
        errs = osqp_errors([o.cpu().numpy() for o in out[:3]], ref[:3])
        assert max(errs.values()) < TOL, (k, errs)
    assert th is not None


def test_full_batch_of_200k(gpu, nine):
    pats, plans, solvers, het = nine
    th = gpu
    counts = hetero_counts(200_000)
    print("hetero full batch:", counts.tolist(), flush=True)
    data = _data(pats, counts, 30)
    sym = []
    for d in data:
        t = lambda v: th.as_tensor(np.ascontiguousarray(np.asarray(v, dtype=np.float32).T), device="cuda")  # noqa: E731
        sym.append(dict(p=t(d[0]), a=t(d[1]), q=t(d[2]), l=t(d[3]), u=t(d[4])))
    outs = het.solve_symbolic(sym)
    th.cuda.synchronize()
    rng = np.random.default_rng(3)
    for k, (plan, d, o) in enumerate(zip(plans, data, outs)):
        for t in o:
            assert bool(th.isfinite(t).all())
        cols = rng.choice(int(counts[k]), size=12, replace=False)
        ref = oracle.OracleOSQP(plan).solve(d[0], d[1], d[2][cols], d[3][cols], d[4][cols])
        for name, t, r in zip(("x", "z", "y"), o[:3], ref[:3]):
            e = rel_err(t[:, cols].t().cpu().numpy(), r)
            assert e < TOL, (k, name, e)
