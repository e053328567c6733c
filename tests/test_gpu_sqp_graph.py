"""One SQP iteration captured in a HIP graph (CompiledSparseMPCSQP.capture_step).

The graph holds linearization (torch.func), the warm-started HIP OSQP launch,
the candidate evaluations and the filter line search, and feeds its results
back into its inputs.  Checks, on the quadrotor MPC (N = 25):

* every replay equals the eager iteration it stands for (same iterate, same
  step lengths and reasons, same warm start), over three iterations, with
  eager work and host synchronisations between the replays;
* loading new inputs and replaying again reproduces the first run;
* a ragged batch (not a multiple of 64 instances).
"""

from __future__ import annotations

import numpy as np
import pytest

from parity_measure import rel_err
from warpmpc_amd.osqp import OSQPSettings
from warpmpc_amd.sqp import build_sparse_mpc_plan, compile_sparse_mpc_sqp
from warpmpc_amd.sqp.problems import quadrotor_initial_guess_and_params, quadrotor_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def quad_sqp(gpu):
    prob = quadrotor_problem(horizon=25)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, scaling=0, rho_is_vec=True)
    plan = build_sparse_mpc_plan(prob, osqp_settings=s)
    return compile_sparse_mpc_sqp(prob, plan, osqp_settings=s)


@pytest.mark.parametrize("batch", [1000, 4096])
def test_replays_equal_eager_iterations(gpu, quad_sqp, batch):
    th = gpu
    sqp = quad_sqp
    z0, params = quadrotor_initial_guess_and_params(batch, seed=6)
    z0 = th.as_tensor(z0, dtype=th.float32, device="cuda")
    params = th.as_tensor(params, dtype=th.float32, device="cuda")
    print(f"[sqp graph] quadrotor batch {batch}: eager iterations", flush=True)
    eager = []
    z, state = z0, sqp.init_state(batch)
    for _ in range(3):
        res, state = sqp.step(z, params, state=state)
        z = res.z_next
        eager.append((z.cpu().numpy(), res.line_search.step_length.cpu().numpy(),
                      res.line_search.reason.cpu().numpy(), state.y.cpu().numpy()))
    print(f"[sqp graph] quadrotor batch {batch}: capture and replay", flush=True)
    graph = sqp.capture_step(batch)
    assert graph.batch_size == batch
    runs = []
    for run in range(2):
        graph.load(z0, params)
        got = []
        for it in range(3):
            graph.replay()
            th.cuda.synchronize()
            got.append((graph.z.cpu().numpy(), graph.last.line_search.step_length.cpu().numpy(),
                        graph.last.line_search.reason.cpu().numpy(), graph.state[2].cpu().numpy()))
            # eager work between replays
            scratch = th.randn((256, 256), device="cuda")
            (scratch @ scratch).sum().item()
        runs.append(got)
        for it, ((gz, gs, gr, gy), (ez, es, er, ey)) in enumerate(zip(got, eager)):
            ez_err, ey_err = rel_err(gz, ez), rel_err(gy, ey)
            print(f"  run {run} iteration {it}: z vs eager {ez_err:.2e}, warm-start y vs eager {ey_err:.2e}",
                  flush=True)
            assert np.isfinite(gz).all()
            assert ez_err < 1e-5 and ey_err < 1e-5, (run, it, ez_err, ey_err)
            np.testing.assert_array_equal(gs, es)
            np.testing.assert_array_equal(gr, er)
    for (a, *_), (b, *_) in zip(*runs):
        assert rel_err(a, b) < 1e-6


def test_load_validates_shapes(gpu, quad_sqp):
    th = gpu
    graph = quad_sqp.capture_step(64)
    with pytest.raises(RuntimeError):
        graph.load(th.zeros((65, quad_sqp.plan.n_variables), device="cuda"),
                   th.zeros((64, quad_sqp.plan.n_parameters), device="cuda"))
    with pytest.raises(ValueError):
        quad_sqp.capture_step(0)
