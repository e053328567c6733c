"""wmpc_osqp_solve_f32_noalloc (the library's fixed per-GPU scratch) against
wmpc_osqp_solve_f32 with a caller-allocated work buffer, through the C ABI.

Both run the same kernels on the same inputs, so outputs must be bit-identical.
Two calls in a row, the second with a smaller batch, check that the shared
scratch keeps nothing from the previous call that matters; shared and
per-instance P/A cover both scratch layouts; an oversized batch must fail with
a clear error and no launch.
"""

from __future__ import annotations

import numpy as np
import pytest

from warpmpc_amd import _lib
from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan, compile_osqp
from warpmpc_amd.problems import batched_mpc_data, mass_chain_mpc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chain(gpu):
    prob = mass_chain_mpc(horizon=10)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s)
    solver = compile_osqp(plan, np.float32)
    return prob, plan, solver.extras["device_plan"]


def _inputs(th, prob, b, seed, per_instance):
    pv, av, q, l, u = batched_mpc_data(prob, b, seed=seed)
    if per_instance:
        rng = np.random.default_rng(seed)
        pv = pv * rng.uniform(0.5, 2.0, size=(b, pv.shape[1]))
        av = np.broadcast_to(av, (b, av.shape[1]))
    sym = lambda v: th.as_tensor(np.ascontiguousarray(np.asarray(v).T), dtype=th.float32, device="cuda")  # noqa: E731
    return {k: sym(v) for k, v in dict(p=pv, a=av, q=q, l=l, u=u).items()}


def _run(th, lib, dev, plan, t, b, pm_ld, noalloc):
    f32 = dict(dtype=th.float32, device="cuda")
    outs = [th.full((plan.n, b), np.nan, **f32), th.full((plan.m, b), np.nan, **f32),
            th.full((plan.m, b), np.nan, **f32)] + [th.full((b,), np.nan, **f32) for _ in range(3)]
    ptrs = [o.data_ptr() for o in outs]
    stream = th.cuda.current_stream().cuda_stream
    common = (dev.handle, t["p"].data_ptr(), t["a"].data_ptr(), pm_ld, t["q"].data_ptr(), t["l"].data_ptr(),
              t["u"].data_ptr(), None, None, None, None, None, 0)
    if noalloc:
        rc = lib.wmpc_osqp_solve_f32_noalloc(*common, *ptrs, 25, b, b, stream)
    else:
        mode = 2 if pm_ld else 1
        work = th.empty((int(lib.wmpc_osqp_work_floats(dev.handle, mode, b)),), **f32)
        rc = lib.wmpc_osqp_solve_f32(*common, work.data_ptr(), work.numel(), *ptrs, 25, b, b, stream)
    _lib.check(rc, "osqp solve")
    th.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("per_instance", [False, True])
def test_noalloc_matches_allocating_entry_point(gpu, chain, per_instance):
    prob, plan, dev = chain
    th = gpu
    lib = _lib.load()
    for b, seed in ((1000, 1), (333, 2)):  # second call smaller than the first
        t = _inputs(th, prob, b, seed, per_instance)
        pm_ld = b if per_instance else 0
        print(f"noalloc vs work: batch {b}, per-instance P/A {per_instance}", flush=True)
        got = _run(th, lib, dev, plan, t, b, pm_ld, noalloc=True)
        ref = _run(th, lib, dev, plan, t, b, pm_ld, noalloc=False)
        for g, r in zip(got, ref):
            assert np.isfinite(r).all()
            np.testing.assert_array_equal(g, r)


def test_noalloc_rejects_batches_beyond_its_scratch(gpu, chain):
    prob, plan, dev = chain
    th = gpu
    lib = _lib.load()
    cap = int(lib.wmpc_osqp_noalloc_max_floats())
    nk = plan.qdldl_plan.n
    b = cap // (2 * nk) + 64  # needs more than the scratch holds in shared-matrix mode
    assert int(lib.wmpc_osqp_work_floats(dev.handle, 1, b)) > cap
    f32 = dict(dtype=th.float32, device="cuda")
    t = {k: th.zeros((w, b), **f32) for k, w in (("q", plan.n), ("l", plan.m), ("u", plan.m))}
    p = th.as_tensor(plan.p_upper.data, dtype=th.float32, device="cuda")
    a = th.as_tensor(plan.a_matrix.data, dtype=th.float32, device="cuda")
    outs = [th.empty((plan.n, b), **f32), th.empty((plan.m, b), **f32), th.empty((plan.m, b), **f32)] + \
        [th.empty((b,), **f32) for _ in range(3)]
    rc = lib.wmpc_osqp_solve_f32_noalloc(dev.handle, p.data_ptr(), a.data_ptr(), 0, t["q"].data_ptr(),
                                         t["l"].data_ptr(), t["u"].data_ptr(), None, None, None, None, None, 0,
                                         *[o.data_ptr() for o in outs], 25, b, b,
                                         th.cuda.current_stream().cuda_stream)
    assert rc != 0
    assert "no-allocation" in lib.wmpc_last_error().decode()
