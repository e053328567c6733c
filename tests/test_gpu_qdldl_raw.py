"""float64 QDLDL and the level-scheduled solve on the GPU.

* float64 factor / solve / factor_and_solve (compile_qdldl_variant(float64))
  against the FP64 oracle, the same operation order: L, D and x within 1e-12
  relative (fused multiply-adds are the only difference); one factor shared
  by a batch of right-hand sides.
* The level-scheduled solve (level_scheduled_solve=True and the C entry point
  with 1 .. 64 lanes per instance and thresholds 1, 2, 8): bitwise equal to
  the FP32 oracle's serial solve given the same float32 factor, whose
  operation order it keeps; shared and per-instance factors, ragged batches;
  the N = 50 microbench KKT at batch 4096 on a sample.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle
from warpmpc_amd import _lib
from warpmpc_amd._device import ptr, stream_ptr
from warpmpc_amd.problems import kkt_microbench_plan, kkt_microbench_values
from warpmpc_amd.qdldl import compile_qdldl_variant

pytestmark = pytest.mark.gpu


def _rel(v, ref):
    v, ref = np.asarray(v, np.float64), np.asarray(ref, np.float64)
    return float(np.max(np.max(np.abs(v - ref), axis=1) / np.maximum(np.max(np.abs(ref), axis=1), 1e-300)))


@pytest.fixture(scope="module")
def chain10(gpu):
    _, plan = kkt_microbench_plan(horizon=10)
    return plan


@pytest.mark.parametrize("batch", [37, 300])
def test_f64_matches_oracle(gpu, chain10, batch):
    th = gpu
    q = chain10.qdldl_plan
    comp = compile_qdldl_variant(q, np.float64, backend="hip", transpose_work=True)
    assert comp.variant == "hip+transpose+f64"
    vals, rhs = kkt_microbench_values(chain10, batch, seed=batch)
    vals, rhs = vals.astype(np.float64), rhs.astype(np.float64)
    print(f"[f64] mass-chain KKT N=10, batch {batch}", flush=True)
    lx, d, dinv = comp.factor(th.as_tensor(vals, device="cuda"))
    assert lx.dtype == th.float64
    lx64, d64, dinv64 = oracle.qdldl_factor(q, vals.T)
    assert _rel(lx.cpu().numpy().T, lx64) < 1e-12
    assert _rel(d.cpu().numpy().T, d64) < 1e-12
    x = comp.solve(lx, dinv, th.as_tensor(rhs, device="cuda"))
    x64 = oracle.qdldl_solve(q, lx64, dinv64, rhs.T)
    assert _rel(x.cpu().numpy().T, x64) < 1e-12
    xf, lxf, _ = comp.factor_and_solve(th.as_tensor(vals, device="cuda"), th.as_tensor(rhs, device="cuda"))
    np.testing.assert_array_equal(xf.cpu().numpy(), x.cpu().numpy())
    # one factor shared by every right-hand side
    xs = comp.solve(lx[:, :1], dinv[:, :1], th.as_tensor(rhs, device="cuda"))
    xs64 = oracle.qdldl_solve(q, lx64[:1], dinv64[:1], rhs.T)
    assert _rel(xs.cpu().numpy().T, xs64) < 1e-12


@pytest.mark.parametrize("threshold", [1, 2, 8])
@pytest.mark.parametrize("lanes", [1, 8, 64])
def test_level_solve_equals_serial_oracle_bitwise(gpu, chain10, threshold, lanes):
    th = gpu
    lib = _lib.load()
    q = chain10.qdldl_plan
    comp = compile_qdldl_variant(q, np.float32, backend="hip", transpose_work=True)
    dev = comp.extras["device_plan"]
    assert lib.wmpc_qdldl_plan_levels(dev.handle) == 230
    b = 70
    vals, rhs = kkt_microbench_values(chain10, b, seed=5)
    lx, _, dinv = comp.factor(th.as_tensor(vals, device="cuda"))
    r = th.as_tensor(rhs, device="cuda")
    print(f"[levels] N=10 KKT, batch {b}, threshold {threshold}, {lanes} lanes per instance", flush=True)
    for shared in (False, True):
        f_ld = 0 if shared else b
        lxs = lx[:, :1].contiguous() if shared else lx
        dis = dinv[:, :1].contiguous() if shared else dinv
        x = th.empty_like(r)
        _lib.check(lib.wmpc_qdldl_solve_levels_f32(dev.handle, threshold, lanes, ptr(lxs), f_ld, ptr(dis), f_ld,
                                                   ptr(r), ptr(x), b, b, stream_ptr()), "levels")
        ref = oracle.qdldl_solve(q, lxs.cpu().numpy().T.astype(np.float64), dis.cpu().numpy().T.astype(np.float64),
                                 rhs.T.astype(np.float64), dtype=np.float32)
        np.testing.assert_array_equal(x.cpu().numpy().T, ref.astype(np.float32))


def test_level_solve_through_compile_and_microbench_pattern(gpu):
    th = gpu
    _, plan = kkt_microbench_plan(horizon=50)
    q = plan.qdldl_plan
    comp = compile_qdldl_variant(q, np.float32, backend="hip", transpose_work=True, level_scheduled_solve=True,
                                 level_scheduled_solve_threshold=4)
    assert comp.variant == "hip+transpose+levelsolve"
    b = 4096
    vals, rhs = kkt_microbench_values(plan, b, seed=2, device="cuda")
    print(f"[levels] N=50 KKT, batch {b} through compile_qdldl_variant", flush=True)
    x, lx, d = comp.factor_and_solve(vals, rhs)
    cols = np.random.default_rng(3).choice(b, 32, replace=False)
    dinv = (1.0 / d[:, cols]).cpu().numpy().T.astype(np.float64)
    ref = oracle.qdldl_solve(q, lx[:, cols].cpu().numpy().T.astype(np.float64), dinv,
                             rhs[:, cols].cpu().numpy().T.astype(np.float64), dtype=np.float32)
    got = x[:, cols].cpu().numpy().T
    assert _rel(got, ref) < 1e-6
    v64 = vals[:, cols].cpu().numpy().T.astype(np.float64)
    lx64, _, dinv64 = oracle.qdldl_factor(q, v64)
    x64 = oracle.qdldl_solve(q, lx64, dinv64, rhs[:, cols].cpu().numpy().T.astype(np.float64))
    assert _rel(got, x64) < 1e-4
    with pytest.raises(_lib.HipError, match="lanes per instance"):
        _lib.check(_lib.load().wmpc_qdldl_solve_levels_f32(comp.extras["device_plan"].handle, 1, 3, ptr(lx), b,
                                                           ptr(d), b, ptr(rhs), ptr(x), b, b, stream_ptr()), "x")
