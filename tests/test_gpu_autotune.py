"""The launch-configuration autotuner (warpmpc_amd.autotune) on the GPU.

* Shared P/A (mass chain N=10) and per-instance P/A: every timed candidate's
  outputs are bitwise equal to its kernel's default configuration (the tuner
  keeps only those; the generated and the generic kernel agree to 1e-5), the
  chosen configuration is applied to the plan and recorded, and solves after
  tuning equal, to float32 rounding, solves before it (bit for bit when the
  kernel did not change).
* apply_cached re-applies a recorded configuration to a fresh solver.
* A heterogeneous plan refuses to launch once one of its patterns' programs
  were re-derived by an LDS budget change, and works again when re-created.
"""

from __future__ import annotations

import numpy as np
import pytest

from warpmpc_amd import _lib, autotune
from warpmpc_amd.osqp import HeterogeneousOSQP, OSQPSettings, build_osqp_plan, compile_osqp
from warpmpc_amd.problems import batched_mpc_data, hetero_mpc_patterns, mass_chain_mpc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chain(gpu):
    prob = mass_chain_mpc(horizon=10)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s)
    return prob, plan


@pytest.mark.parametrize("per_instance", [False, True])
def test_tuning_keeps_results_bitwise(gpu, chain, per_instance, tmp_path):
    prob, plan = chain
    b = 2048
    solver = compile_osqp(plan, np.float32)
    pv, av, q, l, u = batched_mpc_data(prob, b, seed=3)
    if per_instance:
        rng = np.random.default_rng(0)
        pv = np.broadcast_to(pv, (b, pv.shape[1])) * rng.uniform(0.5, 2.0, (b, pv.shape[1]))
        av = np.broadcast_to(av, (b, av.shape[1])).copy()
    before = [o.cpu().numpy() for o in solver.solve(pv, av, q, l, u)]
    print(f"[autotune] mass chain N=10, batch {b}, per_instance={per_instance}", flush=True)
    res = autotune.autotune_osqp(solver, b, per_instance=per_instance, ipw_choices=(16, 64),
                                 lds_choices=(40 * 1024, 96 * 1024), repeats=1, path=tmp_path / "tune.json")
    print(f"  default {res.default} {res.default_ms:.2f} ms -> best {res.best} {res.best_ms:.2f} ms; "
          f"{len(res.timings)} timed, {len(res.rejected)} rejected", flush=True)
    assert len(res.timings) >= 2
    assert all("differ" not in why for _, why in res.rejected), res.rejected
    assert len(res.timings) == len({tuple(sorted(c.items())) for c, _ in res.timings})
    dev = solver.extras["device_plan"]
    assert dev.kkt.ipw_setting == res.best.ipw
    assert dev.kkt.lds_budget == res.best.lds_budget
    assert dev.specialized == res.best.specialized
    after = [o.cpu().numpy() for o in solver.solve(pv, av, q, l, u)]
    for a, b_ in zip(before, after):
        if res.best.specialized == res.default.specialized:
            np.testing.assert_array_equal(a, b_)
        else:
            np.testing.assert_allclose(b_, a, rtol=0, atol=1e-5 * max(1.0, np.abs(a).max()))
    key = autotune.cache_key(dev.pattern_key(), b, per_instance, autotune._device_name())
    assert autotune.lookup(key, tmp_path / "tune.json") == res.best
    fresh = compile_osqp(plan, np.float32)
    assert autotune.apply_cached(fresh, b, per_instance, path=tmp_path / "tune.json")
    assert fresh.extras["device_plan"].kkt.ipw_setting == res.best.ipw
    again = [o.cpu().numpy() for o in fresh.solve(pv, av, q, l, u)]
    for a, b_ in zip(after, again):
        np.testing.assert_array_equal(a, b_)


def test_hetero_plan_detects_rederived_programs(gpu):
    pats = hetero_mpc_patterns()[:2]
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    solvers = [compile_osqp(build_osqp_plan(p.p_matrix, p.a_matrix, p.l, p.u, s), np.float32) for p in pats]
    data = [batched_mpc_data(p, 10, seed=k) for k, p in enumerate(pats)]
    het = HeterogeneousOSQP(solvers)
    first = [o.cpu().numpy() for o in het.solve(data)[0]]
    kkt = solvers[0].extras["device_plan"].kkt
    kkt.set_lds_budget(kkt.lds_budget)
    with pytest.raises(_lib.HipError, match="re-derived"):
        het.solve(data)
    het2 = HeterogeneousOSQP(solvers)
    again = [o.cpu().numpy() for o in het2.solve(data)[0]]
    for a, b_ in zip(first, again):
        np.testing.assert_array_equal(a, b_)
