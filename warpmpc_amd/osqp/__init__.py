"""Fixed-pattern batched OSQP (ADMM) driving the HIP QDLDL kernels."""

from .planning import build_osqp_plan
from .solver import compile_osqp
from .hetero import HeterogeneousOSQP, ShiftMaps, shift_warm_start
from .layout import batch_from_compiled_qdldl, rhs_for_compiled_qdldl, values_for_compiled_qdldl
from .types import CompiledOSQP, FixedOSQPPlan, OSQPDerivativePlan, OSQPSettings, OSQPWarmStart
from ..plan_io import cached_osqp_plan, load_osqp_plan, save_osqp_plan  # noqa: E402 (needs .types)

__all__ = [
    "CompiledOSQP",
    "HeterogeneousOSQP",
    "ShiftMaps",
    "FixedOSQPPlan",
    "OSQPDerivativePlan",
    "OSQPSettings",
    "OSQPWarmStart",
    "build_osqp_plan",
    "compile_osqp",
    "batch_from_compiled_qdldl",
    "cached_osqp_plan",
    "load_osqp_plan",
    "rhs_for_compiled_qdldl",
    "save_osqp_plan",
    "shift_warm_start",
    "values_for_compiled_qdldl",
]
