"""Fixed-pattern batched LDLᵀ (QDLDL) with a HIP backend for MI355X."""

from .core import compile_qdldl, compile_qdldl_variant
from .planning import build_qdldl_plan
from .types import CompiledQDLDL, QDLDLPlan


def save_qdldl_plan(plan, path):
    """Writes a plan in the versioned on-disk format (see ``warpmpc_amd.plan_io``)."""
    from ..plan_io import save_qdldl_plan as save

    return save(plan, path)


def load_qdldl_plan(path):
    """Reads a plan written by ``save_qdldl_plan``."""
    from ..plan_io import load_qdldl_plan as load

    return load(path)


__all__ = ["CompiledQDLDL", "QDLDLPlan", "build_qdldl_plan", "compile_qdldl", "compile_qdldl_variant",
           "load_qdldl_plan", "save_qdldl_plan"]
