"""Fixed-pattern batched LDLᵀ (QDLDL) with a HIP backend for MI355X."""

from .core import compile_qdldl, compile_qdldl_variant
from .planning import build_qdldl_plan
from .types import CompiledQDLDL, QDLDLPlan

__all__ = ["CompiledQDLDL", "QDLDLPlan", "build_qdldl_plan", "compile_qdldl", "compile_qdldl_variant"]
