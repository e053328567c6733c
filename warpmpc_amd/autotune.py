"""Autotuner of the HIP backend's launch configuration, per pattern and batch.

The reference's only tuning is the offline segment-budget sweep of its Warp
kernels (benchmarks/nonlinear_mpc/benchmark_warp_segment_sweep.py).  The HIP
kernels have three knobs that change speed but never results:

* instances per wavefront (8, 16, 32, 64): one lane per instance, the other
  lanes of a wavefront repeating its work; fewer instances per wavefront give
  a small batch more wavefronts (``wmpc_qdldl_plan_set_instances_per_wave``);
* the LDS budget of a workgroup (``wmpc_qdldl_plan_set_lds_budget``): how many
  live L entries the factorization keeps on chip instead of re-reading them,
  against how many workgroups share a CU;
* for a matrix shared by the batch, the generated straight-line ADMM kernel
  or the generic record-stream kernel (``DeviceOSQPPlan.specialize``).

``autotune_osqp`` times every combination on the current GPU with the
solver's own entry point on synthetic inputs of the plan, keeps those whose
outputs are bitwise equal to their kernel's default configuration (the
generated and the generic kernel, which order their float operations
differently, must agree to 1e-5 relative), applies the fastest
to the solver's plan and records it in a JSON cache keyed by the pattern key,
batch size, matrix mode and device name; ``apply_cached`` re-applies a
recorded choice without timing.
"""

from __future__ import annotations

import json
import os
import pathlib
import time
from dataclasses import asdict, dataclass

import numpy as np

from . import _lib
from ._device import ptr, torch

IPW_CHOICES = (8, 16, 32, 64)
LDS_CHOICES = (40 * 1024, 64 * 1024, 96 * 1024, 160 * 1024)


@dataclass(frozen=True)
class LaunchConfig:
    ipw: int                 # instances per wavefront, 0 = automatic
    lds_budget: int          # bytes per workgroup
    specialized: bool        # generated straight-line kernel attached


@dataclass(frozen=True)
class TuneResult:
    best: LaunchConfig
    default: LaunchConfig
    best_ms: float
    default_ms: float
    timings: tuple           # ((LaunchConfig as dict, ms), ...) of the equal-output candidates
    rejected: tuple          # configurations whose outputs differed or that failed


def cache_path() -> pathlib.Path:
    env = os.environ.get("WARPMPC_AMD_TUNE_CACHE")
    return pathlib.Path(env) if env else pathlib.Path.home() / ".cache" / "warpmpc_amd" / "autotune.json"


def cache_key(pattern_key: int, batch: int, per_instance: bool, device: str) -> str:
    return f"{pattern_key:016x}/b{int(batch)}/{'per_instance' if per_instance else 'shared'}/{device}"


def _read_cache(path: pathlib.Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def _write_cache(path: pathlib.Path, data: dict) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_name(path.name + f".tmp{os.getpid()}")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    os.replace(tmp, path)


def record(key: str, config: LaunchConfig, ms: float, path: pathlib.Path | None = None) -> None:
    path = path or cache_path()
    data = _read_cache(path)
    data[key] = {"config": asdict(config), "ms": float(ms)}
    _write_cache(path, data)


def lookup(key: str, path: pathlib.Path | None = None) -> LaunchConfig | None:
    entry = _read_cache(path or cache_path()).get(key)
    if not entry:
        return None
    c = entry["config"]
    return LaunchConfig(ipw=int(c["ipw"]), lds_budget=int(c["lds_budget"]), specialized=bool(c["specialized"]))


def _device_name() -> str:
    th = torch()
    return th.cuda.get_device_name(th.cuda.current_device()).replace("/", "_")


def _current(dev) -> LaunchConfig:
    return LaunchConfig(ipw=int(dev.kkt.ipw_setting), lds_budget=int(dev.kkt.lds_budget),
                        specialized=bool(dev.specialized))


def apply(solver, config: LaunchConfig, code_object: bytes | None = None) -> None:
    """Sets a launch configuration on a compiled HIP OSQP solver's plan."""
    dev = solver.extras["device_plan"]
    dev.kkt.set_instances_per_wave(config.ipw)
    dev.kkt.set_lds_budget(config.lds_budget)
    if config.specialized and not dev.specialized:
        if code_object is None:
            from . import codegen

            path = codegen.find_admm_kernel(solver.plan)
            if path is None:
                raise ValueError("no generated kernel is built for this pattern")
            code_object = path.read_bytes()
        dev.specialize(code_object)
    elif not config.specialized and dev.specialized:
        dev.specialize(None)


def apply_cached(solver, batch: int, per_instance: bool = False, path: pathlib.Path | None = None) -> bool:
    """Applies the recorded configuration of this pattern and batch, if any."""
    dev = solver.extras["device_plan"]
    cfg = lookup(cache_key(dev.pattern_key(), batch, per_instance, _device_name()), path)
    if cfg is None:
        return False
    apply(solver, cfg)
    return True


def _inputs(plan, batch: int, per_instance: bool, seed: int = 0):
    """Batch-major synthetic data of the plan: its own P, A (scaled per
    instance when per_instance), bounds, random q."""
    th = torch()
    rng = np.random.default_rng(seed)
    n, m, b = plan.n, plan.m, batch
    f32 = dict(dtype=th.float32, device="cuda")
    pv = np.asarray(plan.p_upper.data, dtype=np.float64)[:, None]
    av = np.asarray(plan.a_matrix.data, dtype=np.float64)[:, None]
    if per_instance:
        pv = pv * rng.uniform(0.5, 2.0, size=(pv.shape[0], b))
        av = np.repeat(av, b, 1)
    lo, up = _bounds(plan)
    q = rng.standard_normal((n, b))
    t = lambda a: th.as_tensor(np.ascontiguousarray(a), **f32)  # noqa: E731
    return (t(pv), t(av), t(q), t(np.repeat(lo[:, None], b, 1)), t(np.repeat(up[:, None], b, 1)))


def _bounds(plan):
    """Bounds that keep each row's type: equality rows 0, lower/upper-only
    rows -1 / 1 on their finite side, two-sided rows [-1, 1]."""
    ct = np.asarray(plan.constr_type)
    lo = np.where(ct == 1, 0.0, -1.0)
    up = np.where(ct == 1, 0.0, 1.0)
    return lo, up


def autotune_osqp(solver, batch: int, *, per_instance: bool = False, max_iter: int | None = None,
                  ipw_choices=IPW_CHOICES, lds_choices=LDS_CHOICES, repeats: int = 3,
                  path: pathlib.Path | None = None, save: bool = True, stream=None) -> TuneResult:
    """Times the launch configurations of ``solver`` (a HIP ``CompiledOSQP``)
    at ``batch`` instances and applies the fastest whose outputs equal, bit
    for bit, those of its kernel's default configuration."""
    th = torch()
    lib = _lib.load()
    dev = solver.extras["device_plan"]
    plan = solver.plan
    n, m, b = plan.n, plan.m, int(batch)
    if b <= 0:
        raise ValueError("batch must be positive")
    iters = int(max_iter if max_iter is not None else solver.settings.max_iter)
    stream = stream or th.cuda.current_stream()
    pv, av, q, lo, up = _inputs(plan, b, per_instance)
    if not per_instance:
        pv, av = pv[:, :1].contiguous(), av[:, :1].contiguous()
    pm_ld = b if per_instance else 0
    mode = 2 if per_instance else 1
    f32 = dict(dtype=th.float32, device="cuda")
    outs = [th.empty((w, b), **f32) for w in (n, m, m)] + [th.empty((b,), **f32) for _ in range(3)]

    def run():
        work = th.empty((int(lib.wmpc_osqp_work_floats(dev.handle, mode, b)),), **f32)
        _lib.check(lib.wmpc_osqp_solve_f32(
            dev.handle, ptr(pv), ptr(av), pm_ld, ptr(q), ptr(lo), ptr(up), None, None, None, None, None, 0,
            ptr(work), work.numel(), *(ptr(o) for o in outs), iters, b, b, int(stream.cuda_stream)),
            "wmpc_osqp_solve_f32")

    def timed() -> float:
        run()
        th.cuda.synchronize()
        best = float("inf")
        for _ in range(max(1, repeats)):
            t0 = time.perf_counter()
            run()
            th.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best * 1e3

    from . import codegen

    code = None
    if not per_instance:
        found = codegen.find_admm_kernel(plan)
        code = found.read_bytes() if found is not None else None
    start = _current(dev)
    families = (False, True) if code is not None else (False,)
    default = LaunchConfig(ipw=0, lds_budget=start.lds_budget, specialized=start.specialized and code is not None)
    # one reference per kernel: every configuration of a kernel must give its
    # reference's outputs bit for bit; the two kernels differ only in the
    # order of their floating-point operations, so their references must
    # agree to float32 rounding (1e-5 relative per output)
    refs, timings, rejected = {}, [], []
    for spec in families:
        cfg = LaunchConfig(ipw=0, lds_budget=start.lds_budget, specialized=spec)
        apply(solver, cfg, code)
        ms = timed()
        refs[spec] = [o.clone() for o in outs]
        timings.append((asdict(cfg), ms))
    if len(refs) == 2:
        for a, r in zip(refs[True], refs[False]):
            scale = th.clamp_min(r.abs().amax(), 1.0)
            if not bool(((a - r).abs().amax() <= 1e-5 * scale).item()):
                rejected.append(({"specialized": True}, "generated kernel differs from the generic kernel"))
                families = (False,)
                default = LaunchConfig(ipw=0, lds_budget=start.lds_budget, specialized=False)
                timings = [t for t in timings if not t[0]["specialized"]]
                break
    default_ms = next(ms for c, ms in timings if c["specialized"] == default.specialized)
    best, best_ms = default, default_ms
    for c, ms in timings:
        if ms < best_ms:
            best, best_ms = LaunchConfig(**c), ms
    for spec in families:
        for ipw in ipw_choices:
            for lds in lds_choices:
                cfg = LaunchConfig(ipw=int(ipw), lds_budget=int(lds), specialized=spec)
                try:
                    apply(solver, cfg, code)
                    ms = timed()
                except _lib.HipError as exc:  # e.g. an LDS budget below the pattern's fixed windows
                    rejected.append((asdict(cfg), str(exc)))
                    continue
                if not all(th.equal(o, r) for o, r in zip(outs, refs[spec])):
                    rejected.append((asdict(cfg), "outputs differ from the same kernel's default configuration"))
                    continue
                timings.append((asdict(cfg), ms))
                if ms < best_ms:
                    best, best_ms = cfg, ms
    apply(solver, best, code)
    if save:
        record(cache_key(dev.pattern_key(), b, per_instance, _device_name()), best, best_ms, path)
    return TuneResult(best=best, default=default, best_ms=best_ms, default_ms=default_ms, timings=tuple(timings),
                      rejected=tuple(rejected))


__all__ = ["IPW_CHOICES", "LDS_CHOICES", "LaunchConfig", "TuneResult", "apply", "apply_cached", "autotune_osqp",
           "cache_key", "cache_path", "lookup", "record"]
