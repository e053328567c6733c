#!/usr/bin/env python3
"""Benchmark of the batched fixed-pattern QP core on one MI355X.

Headline (``value``): plain forward linear-MPC QP solves/s.  Workload: the
mass-chain MPC (6 unit masses, nx = 12, nu = 4, springs k = 1, ZOH dt = 0.5,
horizon N = 10, Q = I, R = I, |x| <= 4, |u| <= 0.5; n = 172 variables, m = 304
constraints), P and A shared by the batch, x0 ~ U[-1, 1]^12 per instance
(seed 0), batch 65,536, float32, OSQP rho = 0.1 (vector rho), sigma = 1e-6,
alpha = 1.6, 25 ADMM iterations, cold start.  One step = one
``wmpc_osqp_solve_f32`` launch set (the shared KKT factor, then the fused
ADMM kernel) over the whole batch, with inputs resident in HBM in the
library's batch-major layout.

Extra legs on the same JSON line:
* ``fwd_adjoint``: forward + adjoint QP solves/s on the cart-pole MPC
  (N = 30, |u| <= 10, Q = diag(exp theta_q), R = exp theta_r per instance,
  batch 16,384): the forward solve, the loss cotangent, and one fused launch
  of the KKT-form adjoint (the active-set limit of the reference's
  refinement: per-instance equilibration, one factorization on the forward
  KKT pattern, 4 refinement steps), with its error against the FP64 limit;
* ``fwd_adjoint_reference_algorithm``: the same with the reference's own
  adjoint (2*half-square matrix, 100 refinement steps) and its error against
  the FP64 run of the same algorithm;
* ``hetero``: 200,000 instances over the nine mass-chain patterns
  ((nx, nu) in {(4,2), (12,4), (24,8)} x N in {10, 20, 40}) in one
  persistent launch, warm-started from the previous solution shifted one
  stage;
* ``sqp``: the nonlinear quadrotor MPC (N = 25, batch 100,000): 10 SQP
  iterations with the filter line search, 25 ADMM iterations per QP;
* ``qdldl``: the QDLDL factor + solve microbenchmark (mass chain N = 50,
  per-instance KKT values, batch 65,536);
* ``osqp_per_instance``: the headline QP with per-instance P and A, so the
  factor is computed and streamed per instance;
* ``err_vs_fp64`` of each leg against the FP64 oracle on sampled instances;
* ``cpu_baseline``: the oracle's float32 OSQP (OpenMP over instances) on a
  bounded sample of the headline workload, all host cores and one core.

The oracle runs only on the CPU side, after every GPU leg has been timed: the
GPU legs queue their accuracy checks with host copies of their sampled
inputs and outputs, and the CPU-baseline leg runs them before it times the
oracle itself.
"""

from __future__ import annotations

import argparse
import json
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "QP solves/s (fwd; fwd+adjoint) at batch 64k, err≤1e-4 vs FP64, % HBM peak"
S = 4  # bytes per float32


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--legs", default="osqp,adjoint,hetero,sqp,qdldl,per_instance,cpu",
                    help="comma-separated subset of osqp,adjoint,hetero,sqp,qdldl,per_instance,cpu")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--refine", type=int, default=100, help="refinement steps of the reference adjoint")
    ap.add_argument("--kkt-refine", type=int, default=4, help="refinement steps of the KKT-form adjoint")
    ap.add_argument("--ipw", type=int, default=0,
                    help="instances per wavefront of every kernel (0 = the library's automatic choice)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference for the HBM byte counters)")
    return ap.parse_args()


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _copy_calibration(steps: int) -> int:
    """Plain device-to-device copy of 1 GiB: the known-size kernel that calibrates
    the FETCH_SIZE / WRITE_SIZE counters (reads 1 GiB, writes 1 GiB per launch)."""
    import torch

    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_launch": src.numel() * 4, "launches": steps}), flush=True)
    return 0


def _profile_value(name: str, key: str):
    """A number recorded in a committed rocprofv3 summary under profiles/, or None."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    m = re.search(rf"^{key}:\s*([0-9.eE+-]+)", path.read_text(), re.M)
    return float(m.group(1)) if m else None


def _timed(step, steps: int, warmup: int, stream):
    """Warm up, then time `steps` calls with a synchronize on both sides (wall)
    and HIP events on the launch stream around each call (device time)."""
    import numpy as np
    import torch

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    return wall, dev_ms


def _rel_err(v, ref, row_floor=None):
    """Per instance max |v - v64| / max(||v64||_inf, floor): floor 1e-30, or
    ``row_floor`` (one value per instance) for outputs that can be exactly
    zero for an instance."""
    import numpy as np

    v = np.asarray(v, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    floor = 1e-30 if row_floor is None else np.maximum(np.asarray(row_floor, dtype=np.float64), 1e-30)
    return np.max(np.abs(v - ref), axis=1) / np.maximum(np.max(np.abs(ref), axis=1), floor)


def _err_summary(errs: dict) -> dict:
    import numpy as np

    out = {}
    for k, e in errs.items():
        out[k] = {"max": float(np.max(e)), "p999": float(np.percentile(e, 99.9)),
                  "above_1e-4": int(np.sum(e > 1e-4))}
    return out


# Accuracy checks against the FP64 oracle.  The oracle is test infrastructure:
# the GPU legs only queue their checks (with host copies of their sampled
# inputs and outputs), and the CPU-side leg runs them after every GPU leg has
# been timed, next to the CPU baseline.  Each check fills dicts that the JSON
# line already holds.
_CHECKS: list = []


def _check_later(fn) -> None:
    """Queue fn(oracle_module, sqp_oracle_module)."""
    _CHECKS.append(fn)


def _run_checks() -> None:
    from oracle import oracle, sqp_oracle

    _log(f"accuracy checks against the oracle: {len(_CHECKS)}")
    for fn in _CHECKS:
        fn(oracle, sqp_oracle)
    _CHECKS.clear()


def _sym(t):
    """(B, w) batch-first -> contiguous (w, B) batch-major."""
    return t.t().contiguous()


# --------------------------------------------------------------------------- headline: plain forward QP


def _chain_problem():
    import numpy as np

    from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan
    from warpmpc_amd.problems import mass_chain_mpc

    prob = mass_chain_mpc(horizon=10)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s)
    return np, prob, plan


def _osqp_setup(plan, per_instance: bool):
    from warpmpc_amd.osqp import compile_osqp

    solver = compile_osqp(plan, "float32", qdldl_backend="hip", transpose_work=True)
    return solver, solver.extras["device_plan"]


def _u2_bytes(plan, k: int) -> int:
    """U2 model: read P, A, q, l, u; write L, Dinv; K iterations of both sweeps
    over L plus Dinv; write x, z, y, prim, dual, obj."""
    n, m = plan.n, plan.m
    nl, nk = plan.qdldl_plan.nnz_l, plan.qdldl_plan.n
    return S * ((plan.nnz_p + plan.nnz_a) + (n + 2 * m) + (nl + nk) + k * (2 * nl + nk) + (n + 2 * m + 3))


def _osqp_leg(args, per_instance: bool, stream):
    import torch

    from warpmpc_amd import _lib
    from warpmpc_amd.problems import batched_mpc_data

    np, prob, plan = _chain_problem()
    b = args.batch
    _log(f"osqp forward: mass chain N=10 (n={plan.n}, m={plan.m}), batch {b}, 25 iterations, "
         f"{'per-instance' if per_instance else 'shared'} P/A")
    solver, dev = _osqp_setup(plan, per_instance)
    lib = _lib.load()
    pv, av, q, l, u = batched_mpc_data(prob, b, seed=0, dtype=np.float64)
    if per_instance:
        rng = np.random.default_rng(10)
        pv = np.broadcast_to(pv, (b, pv.shape[1])) * rng.uniform(0.5, 2.0, size=(b, pv.shape[1]))
        av = np.broadcast_to(av, (b, av.shape[1])).copy()
    f32 = dict(dtype=torch.float32, device="cuda")
    t = {k: _sym(torch.as_tensor(v, **f32)) for k, v in dict(p=pv, a=av, q=q, l=l, u=u).items()}
    pm_ld = b if per_instance else 0
    n, m = plan.n, plan.m
    x = torch.empty((n, b), **f32)
    z = torch.empty((m, b), **f32)
    y = torch.empty((m, b), **f32)
    prim, dual, obj = (torch.empty((b,), **f32) for _ in range(3))
    mode = 2 if per_instance else 1
    work = torch.empty((int(lib.wmpc_osqp_work_floats(dev.handle, mode, b)),), **f32)
    ptr = lambda v: None if v is None else v.data_ptr()  # noqa: E731

    def step():
        _lib.check(lib.wmpc_osqp_solve_f32(
            dev.handle, ptr(t["p"]), ptr(t["a"]), pm_ld, ptr(t["q"]), ptr(t["l"]), ptr(t["u"]), None, None, None,
            None, None, 0, ptr(work), work.numel(), ptr(x), ptr(z), ptr(y), ptr(prim), ptr(dual), ptr(obj),
            25, b, b, stream.cuda_stream), "wmpc_osqp_solve_f32")

    wall, dev_ms = _timed(step, args.steps, args.warmup, stream)
    # accuracy of every output against the FP64 oracle on sampled instances
    cols = np.random.default_rng(0).choice(b, size=min(256, b), replace=False)
    p_in = pv[cols] if per_instance else pv
    a_in = av[cols] if per_instance else av
    p32 = np.asarray(p_in, dtype=np.float32).astype(np.float64)
    a32 = np.asarray(a_in, dtype=np.float32).astype(np.float64)
    data = (p32, a32, q[cols].astype(np.float32), l[cols].astype(np.float32), u[cols].astype(np.float32))
    got = [v[:, cols].t().cpu().numpy() for v in (x, z, y)]
    errs = {}

    def check(orc, _):
        ref = orc.OracleOSQP(plan).solve(*data)
        errs.update(_err_summary({k: _rel_err(g, r) for k, g, r in zip(("x", "z", "y"), got, ref[:3])}))

    _check_later(check)
    kernel = f"wmpc_admm_shared_{dev.pattern_key():016x}" if dev.specialized and not per_instance else None
    return dict(plan=plan, wall=wall, dev_ms=dev_ms, value=b * args.steps / wall, errs=errs, kernel=kernel)


# --------------------------------------------------------------------------- fwd + adjoint (cart-pole)


def _adjoint_leg(args, stream, mode="reference"):
    import numpy as np
    import torch

    from warpmpc_amd import _lib
    from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan, compile_osqp
    from warpmpc_amd.problems import OSQP_INFTY, cartpole_dynamics, mpc_qp

    b, n_h = 16384, 30
    refine = args.refine if mode == "reference" else args.kkt_refine
    ad, bd = cartpole_dynamics()
    prob = mpc_qp(ad, bd, n_h, np.ones(4), np.ones(1), x_bound=OSQP_INFTY, u_bound=10.0)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s, derivatives=True)
    solver = compile_osqp(plan, "float32", derivatives=True, derivative_refinement_iters=refine,
                          derivative_mode=mode, derivative_kkt_refinement_iters=refine)
    dev, adev = solver.extras["device_plan"], solver.extras["adjoint_plan"]
    lib = _lib.load()
    dp = plan.derivative_plan
    _log(f"fwd+adjoint ({mode} form): cart-pole N={n_h} (n={plan.n}, m={plan.m}, adjoint dim {dp.adjoint_dim}), "
         f"batch {b}, {refine} refinement steps")
    rng = np.random.default_rng(3)
    tq = rng.normal(0.0, 0.5, size=(b, 4))
    tr = rng.normal(0.0, 0.5, size=(b, 1))
    x0 = rng.uniform(-0.2, 0.2, size=(b, 4))
    pv = np.concatenate([np.tile(np.exp(tq), (1, n_h + 1)), np.tile(np.exp(tr), (1, n_h))], 1)
    av = np.broadcast_to(plan.a_matrix.data, (b, plan.nnz_a))
    l = np.tile(prob.l, (b, 1))
    u = np.tile(prob.u, (b, 1))
    l[:, :4] = -x0
    u[:, :4] = -x0
    n, m = plan.n, plan.m
    ui = (n_h + 1) * 4
    f32 = dict(dtype=torch.float32, device="cuda")
    t = {k: _sym(torch.as_tensor(v, **f32)) for k, v in dict(p=pv, a=av, l=l, u=u).items()}
    qz = torch.zeros((n, b), **f32)
    rr = _sym(torch.as_tensor(np.exp(tr), **f32))[0]
    x = torch.empty((n, b), **f32)
    z = torch.empty((m, b), **f32)
    y = torch.empty((m, b), **f32)
    prim, dual, obj = (torch.empty((b,), **f32) for _ in range(3))
    xbar = torch.zeros((n, b), **f32)
    ybar = torch.zeros((m, b), **f32)
    dq = torch.empty((n, b), **f32)
    dl = torch.empty((m, b), **f32)
    du = torch.empty((m, b), **f32)
    dpm = torch.empty((plan.nnz_p, b), **f32)
    dam = torch.empty((plan.nnz_a, b), **f32)
    fwork = torch.empty((int(lib.wmpc_osqp_work_floats(dev.handle, 2, b)),), **f32)
    awork = torch.empty((int(lib.wmpc_adjoint_work_floats(adev.handle, b)),), **f32)

    def forward():
        _lib.check(lib.wmpc_osqp_solve_f32(
            dev.handle, t["p"].data_ptr(), t["a"].data_ptr(), b, qz.data_ptr(), t["l"].data_ptr(),
            t["u"].data_ptr(), None, None, None, None, None, 0, fwork.data_ptr(), fwork.numel(), x.data_ptr(),
            z.data_ptr(), y.data_ptr(), prim.data_ptr(), dual.data_ptr(), obj.data_ptr(), 25, b, b,
            stream.cuda_stream), "wmpc_osqp_solve_f32")

    def adjoint():
        # loss = sum_b R u0^2 (+ terms without x): its cotangent on x is 2 R u0 at u0
        xbar[ui].copy_(2.0 * rr * x[ui])
        _lib.check(lib.wmpc_osqp_adjoint_f32(
            adev.handle, t["p"].data_ptr(), t["a"].data_ptr(), b, t["l"].data_ptr(), t["u"].data_ptr(),
            x.data_ptr(), y.data_ptr(), xbar.data_ptr(), ybar.data_ptr(), refine, awork.data_ptr(), awork.numel(),
            None, dq.data_ptr(), dl.data_ptr(), du.data_ptr(), dpm.data_ptr(), dam.data_ptr(), b, b,
            stream.cuda_stream), "wmpc_osqp_adjoint_f32")

    def step():
        forward()
        adjoint()

    wall, dev_ms = _timed(step, max(2, args.steps // 2), args.warmup, stream)
    steps = max(2, args.steps // 2)
    _, fwd_ms = _timed(forward, steps, 1, stream)
    _, adj_ms = _timed(adjoint, steps, 1, stream)
    cols = np.random.default_rng(1).choice(b, size=96, replace=False)
    f64 = lambda v: v[:, cols].t().cpu().numpy().astype(np.float64)  # noqa: E731
    args_o = (np.asarray(pv[cols], np.float32).astype(np.float64), plan.a_matrix.data[None],
              np.asarray(l[cols], np.float32).astype(np.float64), np.asarray(u[cols], np.float32).astype(np.float64),
              f64(x), f64(y), f64(xbar), f64(ybar))
    got = [f64(v) for v in (dq, dl, du, dpm, dam)]
    names = ("dq", "dl", "du", "dp", "da")
    u2 = _u2_bytes(plan, 25)
    out = dict(b=b, wall=wall, dev_ms=dev_ms, fwd_ms=fwd_ms, adj_ms=adj_ms, steps=steps, u2=u2, plan=plan, dp=dp,
               refine=refine, mode=mode)
    errs, second = {}, {}
    out["errs"] = errs
    if mode == "reference":
        def check(orc, _):
            # the same adjoint in the reference's algorithm by the oracle, FP64 and FP32
            adj = orc.OracleAdjoint(plan, refine=refine)
            r64 = adj.compute(*args_o)[3:]
            r32 = adj.compute(*args_o, dtype=np.float32)[3:]
            errs.update(_err_summary({k: _rel_err(g, r) for k, g, r in zip(names, got, r64)}))
            second.update(_err_summary({k: _rel_err(a, r) for k, a, r in zip(names, r32, r64)}))

        out["spread"] = second
        # U3: B_qp + s [n + 2 (nnz_P + nnz_A) + (n + 2m) + R (2 nnz_L' + n_K') + F (nnz_K' + nnz_L' + n_K')]
        q2 = dp.qdldl_plan
    else:
        ref_refine = args.refine

        def check(orc, _):
            # the active-set limit of the reference's refinement, FP64 (dense solve of M0)
            ex = orc.adjoint_exact(plan, *args_o, clamp_slacks=True)[3:]
            errs.update(_err_summary({k: _rel_err(g, r) for k, g, r in zip(names, got, ex)}))
            r101 = orc.OracleAdjoint(plan, refine=ref_refine).compute(*args_o)[3:]
            second.update(_err_summary({k: _rel_err(a, r) for k, a, r in zip(names, r101, ex)}))

        out["reference_algorithm_vs_limit"] = second
        q2 = plan.qdldl_plan  # the forward KKT pattern
    _check_later(check)
    adj_bytes = S * (n + 2 * (plan.nnz_p + plan.nnz_a) + (n + 2 * m) + (refine + 1) * (2 * q2.nnz_l + q2.n)
                     + 1 * (q2.nnz_a + q2.nnz_l + q2.n))
    out.update(u3=u2 + adj_bytes, adj_bytes=adj_bytes, q2=q2)
    return out


# --------------------------------------------------------------------------- heterogeneous batch (C2)


def _hetero_leg(args, stream):
    import numpy as np
    import torch

    from warpmpc_amd.osqp import HeterogeneousOSQP, OSQPSettings, ShiftMaps, build_osqp_plan, compile_osqp, \
        shift_warm_start
    from warpmpc_amd.problems import batched_mpc_data, hetero_counts, hetero_mpc_patterns, mpc_shift_maps

    total = 200_000
    pats = hetero_mpc_patterns(seed=2)
    counts = hetero_counts(total, seed=2)
    _log(f"heterogeneous batch: 9 patterns, {total} instances {counts.tolist()}, 25 iterations, warm start = "
         f"previous solution shifted one stage")
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plans = [build_osqp_plan(p.p_matrix, p.a_matrix, p.l, p.u, s) for p in pats]
    het = HeterogeneousOSQP([compile_osqp(pl, "float32") for pl in plans])
    f32 = dict(dtype=torch.float32, device="cuda")
    data, sym, maps = [], [], []
    for k, (p, pl, c) in enumerate(zip(pats, plans, counts)):
        d = batched_mpc_data(p, int(c), seed=2 + k)
        data.append(d)
        t = lambda v: torch.as_tensor(np.ascontiguousarray(np.asarray(v).T), **f32)  # noqa: E731
        sym.append(dict(p=t(d[0]), a=t(d[1]), q=t(d[2]), l=t(d[3]), u=t(d[4])))
        maps.append(ShiftMaps(*mpc_shift_maps(p), pl.n, pl.m))
    work = torch.empty((het.work_floats([int(c) for c in counts]),), **f32)
    state = {"out": het.solve_symbolic(sym, work=work)}  # cold start; the timed steps warm-start
    warm = [tuple(torch.empty_like(t) for t in o[:3]) for o in state["out"]]

    def step():
        for m_, o, w in zip(maps, state["out"], warm):
            shift_warm_start(m_, o[0], o[1], o[2], out=w)
        state["out"] = het.solve_symbolic(sym, warm_starts=warm, work=work)

    wall, dev_ms = _timed(step, args.steps, args.warmup, stream)
    # accuracy: sampled instances of the last step against the oracle from the same warm start
    rng = np.random.default_rng(2)
    samples = []
    for pl, d, o, w, c in zip(plans, data, state["out"], warm, counts):
        cols = rng.choice(int(c), size=16, replace=False)
        ws = [t[:, cols].t().cpu().numpy().astype(np.float64) for t in w]
        got = [t[:, cols].t().cpu().numpy() for t in o[:3]]
        samples.append((pl, (d[0], d[1], d[2][cols], d[3][cols], d[4][cols]), ws, got))
    errs = {}

    def check(orc, _):
        acc = {"x": [], "z": [], "y": []}
        for pl, d, ws, got in samples:
            ref = orc.OracleOSQP(pl).solve(*d, x0=ws[0], z0=ws[1], y0=ws[2])
            for name, g, r in zip(("x", "z", "y"), got, ref[:3]):
                acc[name].append(_rel_err(g, r))
        errs.update(_err_summary({k: np.concatenate(v) for k, v in acc.items()}))

    _check_later(check)
    # U4: sum over patterns of U2 (plus the warm start's x0, z0, y0 reads) per instance
    u4 = sum(int(c) * (_u2_bytes(pl, 25) + S * (pl.n + 2 * pl.m)) for pl, c in zip(plans, counts))
    ach = u4 / (dev_ms * 1e-3) / 1e9
    return {"value": total * args.steps / wall, "unit": "QP solves/s", "instances": total,
            "ms_per_step": wall * 1e3 / args.steps, "counts": counts.tolist(),
            "config": "9 mass-chain patterns (nx,nu) in {(4,2),(12,4),(24,8)} x N in {10,20,40}, per-pattern "
                      "springs/Q/R (seed 2), shared P/A per pattern, factor per instance, 25 iterations, warm "
                      "start = previous solution shifted one stage (gather launches inside the step)",
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "model_bytes_per_step": u4, "kernel": "k_osqp_hetero",
                         "traffic": _profile_value("osqp_hetero_nine_patterns_b200k_pmc.txt",
                                                   "traffic_bytes_per_launch")},
            "err_vs_fp64": errs}


# --------------------------------------------------------------------------- nonlinear SQP (C4)


def _sqp_leg(args, stream):
    import numpy as np
    import torch

    from warpmpc_amd.osqp import OSQPSettings
    from warpmpc_amd.sqp import build_sparse_mpc_plan, compile_sparse_mpc_sqp
    from warpmpc_amd.sqp.problems import QUAD_NU, QUAD_NX, quadrotor_initial_guess_and_params, quadrotor_problem

    b, iters = 100_000, 10
    prob = quadrotor_problem(horizon=25)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, scaling=0, rho_is_vec=True)
    plan = build_sparse_mpc_plan(prob, osqp_settings=s, derivatives=True)
    sqp = compile_sparse_mpc_sqp(prob, plan, osqp_settings=s, derivatives=True, derivative_mode="kkt")
    op = plan.osqp_plan
    _log(f"nonlinear SQP: quadrotor N=25 (n={op.n}, m={op.m}, nnz_L={op.qdldl_plan.nnz_l}), batch {b}, "
         f"{iters} line-search SQP iterations x 25 ADMM iterations, one HIP graph per iteration")
    z0, params = quadrotor_initial_guess_and_params(b, seed=4)
    z0 = torch.as_tensor(z0, dtype=torch.float32, device="cuda")
    params = torch.as_tensor(params, dtype=torch.float32, device="cuda")
    graph = sqp.capture_step(b)
    phases = {"linearize": [], "qp": [], "line_search": []}

    def solve_graph():
        graph.load(z0, params)
        return graph.replay(iters)

    def solve_eager(record=False):
        z, state = z0, sqp.init_state(b)
        for _ in range(iters):
            if not record:
                res, state = sqp.step(z, params, state=state)
            else:
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
                ev[0].record(stream)
                lin = sqp.linearize_nlp(z, params)
                ev[1].record(stream)
                solve, state = sqp.solve_qp_data(lin, state=state)
                ev[2].record(stream)
                res = sqp.line_search_step(z, params, solve)
                ev[3].record(stream)
                torch.cuda.synchronize()
                for k, (e0, e1) in zip(phases, zip(ev[:-1], ev[1:])):
                    phases[k].append(e0.elapsed_time(e1))
            z = res.z_next
        return z

    solves = max(1, args.steps // 5)
    wall, _ = _timed(solve_graph, solves, 1, stream)
    wall_eager, _ = _timed(solve_eager, solves, 1, stream)
    solve_eager(record=True)
    z_fin = solve_graph().clone()
    same_as_eager = bool(torch.equal(z_fin, solve_eager()))
    # dataset labels: du0/dx0 through the adjoint of the QP at the final iterate
    # (x0 enters the first stage's rows x - x0 = 0 through l = u = x0 - x)
    lin = sqp.linearize_nlp(z_fin, params)
    fused = sqp.osqp.extras["adjoint_fused"]
    warm = tuple(t.clone() for t in graph.state)
    qp_out = {}

    def labels():
        d, _, y, prim, dual, _ = sqp.osqp.solve_warm_start(lin.p_values, lin.a_values, lin.q, lin.l, lin.u, *warm)
        qp_out.update(d=d, y=y, prim=prim, dual=dual)
        jac = torch.empty((b, QUAD_NU, QUAD_NX), dtype=torch.float32, device="cuda")
        for k in range(QUAD_NU):
            dx = torch.zeros_like(d)
            dx[:, QUAD_NX + k] = 1.0
            _, dl, du, *_ = fused(lin.p_values, lin.a_values, lin.l, lin.u, d, y, dx, None, want_mat=False)
            jac[:, k] = dl[:, :QUAD_NX] + du[:, :QUAD_NX]
        return jac

    jac = labels()
    _, label_ms = _timed(labels, max(1, args.steps // 5), 1, stream)
    # the labels against the FP64 oracle of the same derivative (active-set KKT
    # system at the same QP iterate) on sampled instances.  An instance with
    # every first-stage thrust at a bound has du0/dx0 = 0 exactly (float64 gives
    # ~1e-16, float32 ~1e-10), so each instance is measured against 1 % of its
    # largest adjoint output (the bound rows' own derivative is 1 there) rather
    # than against its Jacobian's norm.
    cols = np.random.default_rng(4).choice(b, size=128, replace=False)
    f64 = lambda t: t[cols].cpu().numpy().astype(np.float64)  # noqa: E731
    lab_in = (f64(lin.p_values), f64(lin.a_values), f64(lin.l), f64(lin.u), f64(qp_out["d"]), f64(qp_out["y"]))
    jac_got = jac[cols].cpu().numpy().reshape(cols.size, -1)
    # one line-search step on a few instances, for the FP64 SQP oracle
    zs, ps = z0[:4].cpu().numpy(), params[:4].cpu().numpy()
    res, _ = sqp.step(z0[:4], params[:4], state=sqp.init_state(4))
    step_dir = res.solve.direction.cpu().numpy()
    step_len = res.line_search.step_length.cpu().numpy()
    lab_err, first_step, lab_zero = {}, {}, {}

    def check(orc, sqp_orc):
        ref = np.zeros((cols.size, QUAD_NU, QUAD_NX))
        grad_scale = np.zeros(cols.size)
        for k in range(QUAD_NU):
            dx = np.zeros((cols.size, op.n))
            dx[:, QUAD_NX + k] = 1.0
            ex = orc.adjoint_exact(op, *lab_in, dx, np.zeros((cols.size, op.m)), clamp_slacks=True)
            ref[:, k] = ex[4][:, :QUAD_NX] + ex[5][:, :QUAD_NX]
            for r in ex[3:6]:
                grad_scale = np.maximum(grad_scale, np.abs(r).reshape(cols.size, -1).max(1))
        jac_err = _rel_err(jac_got, ref.reshape(cols.size, -1), row_floor=1e-2 * grad_scale)
        lab_err.update(_err_summary({"du0_dx0": jac_err}))
        lab_zero["sampled_with_zero_jacobian"] = int(np.sum(np.abs(ref).reshape(cols.size, -1).max(1) < 1e-12))
        sref = sqp_orc.OracleSQP(prob, plan).step(zs, ps)
        first_step.update({"direction": float(np.max(_rel_err(step_dir, sref["direction"]))),
                           "same_step_lengths": bool(np.array_equal(step_len, sref["step_length"]))})

    _check_later(check)
    ms = wall * 1e3 / solves
    qp_ms = float(np.mean(phases["qp"]))
    u2 = _u2_bytes(op, 25) + S * (op.n + 2 * op.m)  # warm-started QP
    ach = u2 * b / (qp_ms * 1e-3) / 1e9
    return {"value": b * solves / wall, "unit": "SQP solves/s (10 iterations each)", "qp_solves_per_s":
            b * iters * solves / wall, "batch": b, "ms_per_solve": ms, "ms_per_solve_eager": wall_eager * 1e3 / solves,
            "graph_equals_eager": same_as_eager,
            "phase_ms_per_iteration_eager": {k: float(np.mean(v)) for k, v in phases.items()},
            "finite": bool(torch.isfinite(z_fin).all()),
            "labels_du0_dx0": {"ms": label_ms, "per_s": b / (label_ms * 1e-3),
                               "method": "QP at the final iterate (25 ADMM iterations, warm start) + 4 KKT-form "
                                         "adjoint launches (one per input): du0/dx0 of the active-set KKT system",
                               "err_vs_fp64": lab_err,
                               "err_measure": "per instance |J - J64|_inf / max(|J64|_inf, 1e-2 x the instance's "
                                              "largest adjoint output), 128 sampled instances",
                               "zero_jacobian": lab_zero,
                               "qp_prim_res_median": float(qp_out["prim"].median()),
                               "qp_dual_res_median": float(qp_out["dual"].median())},
            "config": "quadrotor (12 states, 4 thrusts, RK4 dt 0.02) N=25, x0 seed 4, hover initial guess, "
                      "10 SQP iterations with the filter line search, 25 ADMM iterations per QP (warm start from "
                      "the previous QP), each iteration one HIP graph replay; linearization by torch.func on the "
                      "GPU, QP by the HIP OSQP",
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS,
                         "kernel": "QP phase: k_osqp_admm (per-instance factor) + layout copies",
                         "model_bytes_per_qp": u2, "kernel_ms": qp_ms},
            "err_vs_fp64_first_step": first_step}


# --------------------------------------------------------------------------- QDLDL microbench (C1)


def _qdldl_leg(args, stream):
    import numpy as np
    import torch

    from warpmpc_amd import _lib
    from warpmpc_amd.problems import kkt_microbench_plan, kkt_microbench_values
    from warpmpc_amd.qdldl import compile_qdldl_variant

    b = args.batch
    _log(f"qdldl factor+solve: mass-chain KKT N=50, batch {b}")
    _, oplan = kkt_microbench_plan(horizon=50)
    q = oplan.qdldl_plan
    comp = compile_qdldl_variant(q, np.float32, backend="hip", transpose_work=True)
    dev = comp.extras["device_plan"]
    lib = _lib.load()
    vals, rhs = kkt_microbench_values(oplan, b, seed=1, device="cuda")
    f32 = dict(dtype=torch.float32, device="cuda")
    x = torch.empty((q.n, b), **f32)
    lx = torch.empty((q.nnz_l, b), **f32)
    d = torch.empty((q.n, b), **f32)
    dinv = torch.empty((q.n, b), **f32)
    work = torch.empty((q.n, b), **f32)

    def step():
        _lib.check(lib.wmpc_qdldl_factor_solve_f32(dev.handle, vals.data_ptr(), rhs.data_ptr(), x.data_ptr(),
                                                   lx.data_ptr(), d.data_ptr(), dinv.data_ptr(), work.data_ptr(),
                                                   work.numel(), b, b, stream.cuda_stream), "factor_solve")

    wall, dev_ms = _timed(step, args.steps, args.warmup, stream)
    cols = np.random.default_rng(0).choice(b, size=min(256, b), replace=False)
    v64 = vals[:, cols].cpu().numpy().T.astype(np.float64)
    r64 = rhs[:, cols].cpu().numpy().T.astype(np.float64)
    x_got = x[:, cols].cpu().numpy().T
    errs = {}

    def check(orc, _):
        lx64, _, dinv64 = orc.qdldl_factor(q, v64)
        x64 = orc.qdldl_solve(q, lx64, dinv64, r64)
        errs.update(_err_summary({"x": _rel_err(x_got, x64)}))

    _check_later(check)
    # U1: read values + rhs; write L, D, Dinv; re-read L in the backward sweep; write x
    u1 = S * (q.nnz_a + 2 * q.nnz_l + 4 * q.n)
    ach = u1 * b / (dev_ms * 1e-3) / 1e9
    return {"value": b * args.steps / wall, "unit": "factor+solve systems/s", "batch": b, "kernel_ms": dev_ms,
            "config": f"mass-chain MPC KKT N=50 (n_kkt={q.n}, nnz_kkt={q.nnz_a}, nnz_L={q.nnz_l}), "
                      f"per-instance values",
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "bytes_per_system": u1,
                         "traffic": _profile_value("qdldl_factor_solve_n50_b65536_pmc.txt",
                                                   "traffic_bytes_per_launch")},
            "err_vs_fp64": errs}


# --------------------------------------------------------------------------- CPU baseline


def _cpu_model() -> str:
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_leg(args, plan):
    import os

    import numpy as np

    from oracle import oracle
    from warpmpc_amd.problems import batched_mpc_data, mass_chain_mpc

    prob = mass_chain_mpc(horizon=10)
    chunk = 1024
    pv, av, q, l, u = batched_mpc_data(prob, chunk, seed=0, dtype=np.float32)
    orc = oracle.OracleOSQP(plan)
    cores = len(os.sched_getaffinity(0))
    out = {}
    for threads, secs in ((cores, args.cpu_seconds), (1, args.cpu_seconds / 3)):
        oracle.set_threads(threads)
        orc.solve(pv, av, q[:64], l[:64], u[:64], dtype=np.float32)  # warm-up
        _log(f"cpu baseline: float32 oracle OSQP, {threads} thread(s), >= {secs:.0f} s")
        done, reps = 0, []
        t_end = time.perf_counter() + secs
        while True:
            t0 = time.perf_counter()
            orc.solve(pv, av, q, l, u, dtype=np.float32)
            reps.append(time.perf_counter() - t0)
            done += chunk
            if time.perf_counter() >= t_end and len(reps) >= 5:
                break
        out[threads] = (chunk / float(np.median(reps)), done, sum(reps), len(reps))
    oracle.set_threads(cores)
    v_all, done_all, t_all, n_all = out[cores]
    v_one, done_one, t_one, n_one = out[1]
    return {"value": v_all, "unit": "QP solves/s", "cores": cores, "kind": "port",
            "one_core_value": v_one, "cpu_model": _cpu_model(),
            "sample": f"the headline workload (mass chain N=10, 25 ADMM iterations, shared P/A) in chunks of "
                      f"{chunk} instances, value = chunk / median chunk time: {n_all} chunks ({done_all} QPs, "
                      f"{t_all:.1f} s) on {cores} cores, {n_one} chunks ({done_one} QPs, {t_one:.1f} s) on 1 core; "
                      f"float32 restatement of the reference's op order, OpenMP over instances"}


# --------------------------------------------------------------------------- main


def main() -> int:
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (instances shard with no "
              "communication, so N GPUs run N independent copies of this workload)", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return _copy_calibration(args.steps)
    import torch

    legs = set(args.legs.split(","))
    torch.cuda.set_device(0)
    from warpmpc_amd import _lib

    _lib.check(_lib.load().wmpc_set_instances_per_wave(args.ipw), "wmpc_set_instances_per_wave")
    stream = torch.cuda.current_stream()
    b = args.batch
    line = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ipw": args.ipw,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic"}
    plan = None
    if "osqp" in legs:
        r = _osqp_leg(args, False, stream)
        plan = r["plan"]
        u2 = _u2_bytes(plan, 25)
        ach = u2 * b / (r["dev_ms"] * 1e-3) / 1e9
        traffic = _profile_value("osqp_admm_chain_n10_b65536_pmc.txt", "traffic_bytes_per_launch")
        kname = "k_osqp_factor_shared + " + (r["kernel"] or "k_osqp_admm<true, false>")
        line.update({
            "value": r["value"], "unit": "QP solves/s", "ms_per_step": r["wall"] * 1e3 / args.steps,
            "config": {"workload": f"linear-MPC QP, mass chain nx=12 nu=4 N=10 (n={plan.n}, m={plan.m}, "
                                   f"nnz_L={plan.qdldl_plan.nnz_l}), shared P/A, x0~U[-1,1]^12 seed 0, "
                                   f"25 ADMM iterations, cold start",
                       "batch": b, "max_iter": 25, "parallelism": "single GPU, one lane per instance"},
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "model_bytes_per_qp": u2,
                         "kernel": kname, "kernel_ms": r["dev_ms"],
                         "measured_GBps": None if traffic is None else traffic / (r["dev_ms"] * 1e-3) / 1e9,
                         "note": "achieved = U2 model bytes (L and Dinv streamed per instance) / kernel time; "
                                 "with P/A shared by the batch the factor is one wave-uniform, cache-resident "
                                 "stream, so HBM sees only the per-instance iterate (traffic = measured bytes "
                                 "per launch, measured_GBps = traffic / kernel time); the kernel is bound by "
                                 "operand latency at one wavefront per SIMD"},
            "err_vs_fp64": r["errs"],
        })
    for leg, mode in (("adjoint", "kkt"), ("adjoint_reference", "reference")):
        if leg not in legs and not (leg == "adjoint_reference" and "adjoint" in legs):
            continue
        r = _adjoint_leg(args, stream, mode)
        # dominant kernel: the adjoint launch, with the adjoint part of U3
        ach = r["adj_bytes"] * r["b"] / (r["adj_ms"] * 1e-3) / 1e9
        q2 = r["q2"]
        kkt = mode == "kkt"
        prof = "osqp_adjoint_kkt_cartpole_n30_b16384_pmc.txt" if kkt else "osqp_adjoint_cartpole_n30_b16384_pmc.txt"
        entry = {
            "value": r["b"] * r["steps"] / r["wall"], "unit": "fwd+adjoint QP solves/s", "batch": r["b"],
            "ms_per_step": r["wall"] * 1e3 / r["steps"], "forward_ms": r["fwd_ms"], "adjoint_ms": r["adj_ms"],
            "fwd_only_qp_per_s": r["b"] / (r["fwd_ms"] * 1e-3),
            "R": r["refine"] + 1, "F": 1,
            "config": (f"cart-pole N=30 (n={r['plan'].n}, m={r['plan'].m}), per-instance Q/R, 25 ADMM iterations; "
                       + (f"KKT-form adjoint: one factorization of the forward KKT pattern (dim {q2.n}, nnz_L "
                          f"{q2.nnz_l}) and {r['refine']} refinement steps against the exact system"
                          if kkt else
                          f"reference adjoint: matrix of dim {q2.n} (nnz {q2.nnz_a}, nnz_L {q2.nnz_l}) factored "
                          f"per instance, {r['refine']} refinement steps")),
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "kernel": "k_osqp_adjoint_kkt" if kkt else "k_osqp_adjoint",
                         "kernel_ms": r["adj_ms"], "model_bytes_per_instance": r["adj_bytes"],
                         "u3_bytes_per_qp": r["u3"],
                         "traffic": _profile_value(prof, "traffic_bytes_per_launch")},
            "err_vs_fp64": r["errs"],
        }
        if kkt:
            entry["err_reference"] = "FP64 limit of the reference's refinement (M0 solved exactly)"
            entry["reference_algorithm_vs_limit"] = r["reference_algorithm_vs_limit"]
            line["fwd_adjoint"] = entry
        else:
            entry["err_reference"] = "FP64 run of the reference's algorithm with the same R"
            entry["reference_fp32_spread"] = r["spread"]
            line["fwd_adjoint_reference_algorithm"] = entry
    if "hetero" in legs:
        line["hetero"] = _hetero_leg(args, stream)
    if "sqp" in legs:
        line["sqp"] = _sqp_leg(args, stream)
    if "qdldl" in legs:
        line["qdldl"] = _qdldl_leg(args, stream)
    if "per_instance" in legs:
        r = _osqp_leg(args, True, stream)
        u2 = _u2_bytes(r["plan"], 25)
        ach = u2 * b / (r["dev_ms"] * 1e-3) / 1e9
        line["osqp_per_instance"] = {
            "value": r["value"], "unit": "QP solves/s", "batch": b, "kernel_ms": r["dev_ms"],
            "config": "headline QP with per-instance P (x U[0.5,2]) and A: factor per instance",
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "model_bytes_per_qp": u2,
                         "kernel": "k_osqp_admm<false, true>", "kernel_ms": r["dev_ms"],
                         "traffic": _profile_value("osqp_admm_per_instance_chain_n10_b65536_pmc.txt",
                                                   "traffic_bytes_per_launch")},
            "err_vs_fp64": r["errs"],
        }
    # the CPU side: accuracy checks of the legs above, then the CPU baseline
    _run_checks()
    if "cpu" in legs:
        if plan is None:
            _, _, plan = _chain_problem()
        line["cpu_baseline"] = _cpu_leg(args, plan)
    if "value" not in line and "fwd_adjoint" in line:
        line.update(value=line["fwd_adjoint"]["value"], unit=line["fwd_adjoint"]["unit"],
                    ms_per_step=line["fwd_adjoint"]["ms_per_step"], config={"workload": "fwd+adjoint only"})
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
