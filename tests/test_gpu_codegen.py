"""Generated straight-line ADMM kernels (warpmpc_amd/codegen.py) on the GPU.

The generated kernel of a pattern must give the generic record-stream
kernel's results on the same inputs and the FP64 oracle's to the usual 1e-4.
Both are float32 and accumulate the backward sums in different groupings, so
they agree to rounding, not bitwise: each lies about 5e-6 from FP64 on these
cases, and their mutual difference is held to 5e-5 for x, z, y and 1e-4 for
the residuals (maxima of cancelling sums) and the objective.  Cases: the
mass-chain MPC (N = 10) with shared P/A: cold and warm start, ragged batches
(1, 65, 1000),
max_iter 0 and 1, a passed shared factor; the library's pattern key equals
the generator's; a code object of another pattern is refused.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from parity_measure import osqp_errors
from warpmpc_amd import codegen
from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan, compile_osqp
from warpmpc_amd.problems import batched_mpc_data, mass_chain_mpc

pytestmark = pytest.mark.gpu


def _plan(horizon: int, max_iter: int = 25):
    prob = mass_chain_mpc(horizon=horizon)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=max_iter, rho_is_vec=True)
    return prob, build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s)


@pytest.fixture(scope="module")
def pair(gpu):
    prob, plan = _plan(10)
    if codegen.find_admm_kernel(plan) is None:
        pytest.skip("no generated kernel built for the N=10 chain (run warpmpc_amd.build)")
    gen = compile_osqp(plan, np.float32, specialize="auto")
    ref = compile_osqp(plan, np.float32, specialize=False)
    assert gen.extras["device_plan"].specialized
    assert not ref.extras["device_plan"].specialized
    return prob, plan, gen, ref


def _compare(out_g, out_r, q, ref64=None, tol_vec=5e-5, tol_scalar=1e-4, tol64=1e-4):
    """Generated against generic kernel, and against the FP64 oracle, by the
    parity tests' measure (norm-wise relative per instance)."""
    errs = osqp_errors([o.cpu().numpy() for o in out_g], [o.cpu().numpy() for o in out_r], q)
    for name, e in errs.items():
        tol = tol_vec if name in ("x", "z", "y") else tol_scalar
        assert e < tol, f"{name}: generated vs generic {e}"
    if ref64 is not None:
        errs = osqp_errors([o.cpu().numpy() for o in out_g], ref64, q)
        for name, e in errs.items():
            assert e < tol64, f"{name}: generated vs FP64 {e}"


def test_pattern_key_matches_generator(pair):
    _, plan, gen, _ = pair
    assert gen.extras["device_plan"].pattern_key() == codegen.plan_key(plan)


@pytest.mark.parametrize("batch", [1, 65, 1000])
def test_cold_start_matches_generic_and_oracle(pair, batch):
    prob, plan, gen, ref = pair
    pv, av, q, l, u = batched_mpc_data(prob, batch, seed=batch)
    print(f"generated kernel, cold start, batch {batch}", flush=True)
    out_g = gen.solve(pv, av, q, l, u)
    out_r = ref.solve(pv, av, q, l, u)
    ref64 = oracle.OracleOSQP(plan).solve(pv, av, q, l, u)
    _compare(out_g, out_r, q, ref64)


def test_warm_start_matches_generic_and_oracle(pair):
    prob, plan, gen, ref = pair
    b = 300
    pv, av, q, l, u = batched_mpc_data(prob, b, seed=7)
    rng = np.random.default_rng(7)
    ws = (rng.normal(size=(b, plan.n)), rng.normal(size=(b, plan.m)), rng.normal(size=(b, plan.m)))
    print("generated kernel, warm start, batch 300", flush=True)
    out_g = gen.solve_warm_start(pv, av, q, l, u, *ws)
    out_r = ref.solve_warm_start(pv, av, q, l, u, *ws)
    ref64 = oracle.OracleOSQP(plan).solve(pv, av, q, l, u, x0=ws[0], z0=ws[1], y0=ws[2])
    _compare(out_g, out_r, q, ref64)


@pytest.mark.parametrize("max_iter", [0, 1])
def test_few_iterations(gpu, max_iter):
    prob, plan = _plan(10, max_iter=max_iter)
    if codegen.find_admm_kernel(plan) is None:
        pytest.skip("no generated kernel built for the N=10 chain")
    gen = compile_osqp(plan, np.float32, specialize="auto")
    ref = compile_osqp(plan, np.float32, specialize=False)
    pv, av, q, l, u = batched_mpc_data(prob, 130, seed=3)
    print(f"generated kernel, max_iter {max_iter}", flush=True)
    _compare(gen.solve(pv, av, q, l, u), ref.solve(pv, av, q, l, u), q,
             oracle.OracleOSQP(plan, max_iter=max_iter).solve(pv, av, q, l, u))


def test_passed_shared_factor(pair):
    prob, plan, gen, ref = pair
    b = 200
    pv, av, q, l, u = batched_mpc_data(prob, b, seed=11)
    print("generated kernel, passed shared factor", flush=True)
    lx, dinv = gen.factor(pv, av)
    out_g = gen.solve_with_factor(lx, dinv, pv, av, q, l, u)
    out_r = ref.solve_with_factor(lx, dinv, pv, av, q, l, u)
    _compare(out_g, out_r, q, oracle.OracleOSQP(plan).solve(pv, av, q, l, u))


def test_foreign_code_object_is_refused(pair):
    _, _, gen, _ = pair
    _, plan20 = _plan(20)
    other = codegen.find_admm_kernel(plan20)
    if other is None:
        pytest.skip("no generated kernel built for the N=20 chain")
    dev = gen.extras["device_plan"]
    with pytest.raises(RuntimeError, match="another pattern"):
        dev.specialize(other.read_bytes())
    # the refused image detached nothing else: reattach the plan's own kernel
    dev.specialize(codegen.find_admm_kernel(pair[1]).read_bytes())
    assert dev.specialized
