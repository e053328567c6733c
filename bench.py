#!/usr/bin/env python3
"""Benchmark of the batched fixed-pattern QDLDL factor + solve on one MI355X.

Workload (the configuration the headline metric is quoted on): the KKT matrix
[P + sigma I, Aᵀ; A, -diag(1/rho)] of the mass-chain linear MPC (6 unit masses,
nx = 12, nu = 4, springs k = 1, ZOH dt = 0.5) at horizon N = 50, batch 65,536,
float32, per-instance values (P diagonal 1 + U[0,1], rho log-uniform in
[1e-3, 1e2] per constraint), rhs ~ N(0, 1), seed 1.  One step = one fused
factor + solve launch over the whole batch, inputs resident in HBM.

Also reported (extra keys of the same JSON line):
* ``osqp_fwd``: plain forward linear-MPC QP solves/s (mass chain N = 10,
  batch 65,536, 25 ADMM iterations, shared P/A);
* ``err_vs_fp64``: worst-case and 99.9th-percentile per-instance relative error
  of x against the FP64 oracle on a sample of instances;
* ``roofline``: algorithmic HBM bytes per launch / measured kernel time;
* ``cpu_baseline``: the oracle's float32 CPU counterpart (OpenMP) on a bounded
  sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "QP solves/s (fwd; fwd+adjoint) at batch 64k, err≤1e-4 vs FP64, % HBM peak"


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--horizon", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-osqp", action="store_true", help="skip the OSQP forward leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy (reference for the HBM byte counters)")
    return ap.parse_args()


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


def _traffic_from_profile(name: str):
    """HBM bytes per launch recorded by the committed rocprofv3 PMC summary, or None."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", path.read_text(), re.M)
    return float(m.group(1)) if m else None


def main() -> int:
    args = _parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (instances shard with no "
              "communication, so N GPUs run N independent copies of this workload)", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return _copy_calibration(args.steps)
    import numpy as np
    import torch

    from warpmpc_amd import _lib
    from warpmpc_amd.problems import kkt_microbench_plan, kkt_microbench_values
    from warpmpc_amd.qdldl import compile_qdldl_variant

    torch.cuda.set_device(0)
    lib = _lib.load()
    b = args.batch
    print(f"[bench] planning mass-chain KKT N={args.horizon}", file=sys.stderr, flush=True)
    _, oplan = kkt_microbench_plan(horizon=args.horizon)
    q = oplan.qdldl_plan
    comp = compile_qdldl_variant(q, np.float32, backend="hip", transpose_work=True)
    dev = comp.extras["device_plan"]
    vals, rhs = kkt_microbench_values(oplan, b, seed=1, device="cuda")
    f32 = dict(dtype=torch.float32, device="cuda")
    x = torch.empty((q.n, b), **f32)
    lx = torch.empty((q.nnz_l, b), **f32)
    d = torch.empty((q.n, b), **f32)
    dinv = torch.empty((q.n, b), **f32)
    work = torch.empty((q.n, b), **f32)
    stream = torch.cuda.current_stream()

    def step():
        _lib.check(lib.wmpc_qdldl_factor_solve_f32(dev.handle, vals.data_ptr(), rhs.data_ptr(), x.data_ptr(),
                                                   lx.data_ptr(), d.data_ptr(), dinv.data_ptr(), work.data_ptr(),
                                                   work.numel(), b, b, stream.cuda_stream), "factor_solve")

    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    ms_per_step = wall * 1e3 / args.steps
    value = b * args.steps / wall

    # accuracy vs the FP64 oracle on a sample of instances
    from oracle import oracle

    cols = np.random.default_rng(0).choice(b, size=min(256, b), replace=False)
    v64 = vals[:, cols].cpu().numpy().T.astype(np.float64)
    r64 = rhs[:, cols].cpu().numpy().T.astype(np.float64)
    lx64, _, dinv64 = oracle.qdldl_factor(q, v64)
    x64 = oracle.qdldl_solve(q, lx64, dinv64, r64)
    xs = x[:, cols].cpu().numpy().T.astype(np.float64)
    rel = np.max(np.abs(xs - x64), axis=1) / np.max(np.abs(x64), axis=1)

    nnz_k, nnz_l, nk = q.nnz_a, q.nnz_l, q.n
    # values + rhs in; L, D, Dinv out; L re-read by the backward sweep; x out
    bytes_per_sys = 4 * (nnz_k + 2 * nnz_l + 4 * nk)
    achieved = bytes_per_sys * b / (kern_ms * 1e-3) / 1e9
    traffic = _traffic_from_profile("qdldl_factor_solve_n50_b65536_pmc.txt")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "factor+solve systems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"qdldl factor+solve, mass-chain MPC KKT N={args.horizon} (n_kkt={nk}, nnz_kkt={nnz_k}, "
                        f"nnz_L={nnz_l}), per-instance values, batch {b}",
            "batch": b,
            "horizon": args.horizon,
            "n_kkt": nk,
            "nnz_kkt": nnz_k,
            "nnz_l": nnz_l,
            "parallelism": "single GPU, one lane per instance",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_system": bytes_per_sys,
            "kernel_ms": kern_ms,
        },
        "err_vs_fp64": {"max": float(rel.max()), "p999": float(np.percentile(rel, 99.9)),
                        "sample": int(cols.size)},
    }

    if not args.no_osqp:
        line["osqp_fwd"] = _osqp_leg(args)
    if not args.no_cpu:
        line["cpu_baseline"] = _cpu_leg(q, args)
    print(json.dumps(line), flush=True)
    return 0


def _osqp_leg(args):
    import numpy as np
    import torch

    from oracle import oracle
    from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan, compile_osqp
    from warpmpc_amd.problems import batched_mpc_data, mass_chain_mpc

    b = args.batch
    print("[bench] osqp forward leg: mass chain N=10, 25 iterations", file=sys.stderr, flush=True)
    prob = mass_chain_mpc(horizon=10)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s)
    solver = compile_osqp(plan, np.float32, qdldl_backend="hip", transpose_work=True)
    pv, av, q, l, u = batched_mpc_data(prob, b, seed=0, dtype=np.float32)
    t = [torch.as_tensor(a).cuda() for a in (pv, av, q, l, u)]
    out = solver.solve(*t)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        out = solver.solve(*t)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    cols = np.random.default_rng(0).choice(b, size=128, replace=False)
    ref = oracle.OracleOSQP(plan).solve(pv, av, q[cols], l[cols], u[cols])
    xs = out[0].cpu().numpy()[cols].astype(np.float64)
    rel = np.max(np.abs(xs - ref[0]), axis=1) / np.maximum(np.max(np.abs(ref[0]), axis=1), 1e-30)
    return {"value": b / dt, "unit": "QP solves/s", "batch": b, "ms": dt * 1e3, "max_iter": 25,
            "config": "mass chain N=10 (n=172, m=304), shared P/A, rho_is_vec",
            "err_vs_fp64": {"max": float(rel.max()), "p999": float(np.percentile(rel, 99.9))}}


def _cpu_leg(q, args):
    import numpy as np

    from oracle import oracle
    from warpmpc_amd.problems import kkt_microbench_plan, kkt_microbench_values

    _, oplan = kkt_microbench_plan(horizon=args.horizon)
    chunk = 2048
    vals, rhs = kkt_microbench_values(oplan, chunk, seed=2)
    v = np.ascontiguousarray(vals.T)
    r = np.ascontiguousarray(rhs.T)
    threads = oracle.num_threads()
    print(f"[bench] cpu baseline: float32 oracle, {threads} threads, >= {args.cpu_seconds:.0f} s",
          file=sys.stderr, flush=True)
    done = 0
    t0 = time.perf_counter()
    while True:
        lx, _, dinv = oracle.qdldl_factor(q, v, dtype=np.float32)
        oracle.qdldl_solve(q, lx, dinv, r, dtype=np.float32)
        done += chunk
        if time.perf_counter() - t0 >= args.cpu_seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "factor+solve systems/s", "cores": threads, "kind": "port",
            "sample": f"{done} systems of the same N={args.horizon} KKT workload in chunks of {chunk}, "
                      f"float32, OpenMP over instances, {dt:.1f} s"}


if __name__ == "__main__":
    sys.exit(main())
