"""Fixed-pattern batched OSQP (ADMM) driving the HIP QDLDL kernels."""

from .planning import build_osqp_plan
from .solver import compile_osqp
from .hetero import HeterogeneousOSQP, ShiftMaps, shift_warm_start
from .types import CompiledOSQP, FixedOSQPPlan, OSQPDerivativePlan, OSQPSettings, OSQPWarmStart

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
    "shift_warm_start",
]
