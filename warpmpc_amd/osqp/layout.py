"""Layout adapters between the batch-first OSQP arrays and a compiled QDLDL.

Same names and semantics as ``warpmpc/jax_osqp/solver.py:665-682``: a
``CompiledQDLDL`` declares the layout its callables take
(``values_layout``: ``"original_batch"`` = (B, nnz) or ``"original_symbolic"``
= (nnz, B); ``rhs_layout``: ``"batch"`` = (B, n) or ``"symbolic"`` = (n, B)),
and these helpers move a batch-first array into it and back.  The HIP backend
always uses the symbolic (batch-major) layouts.  Arrays may be numpy arrays or
torch tensors; the result is a transposed view, as ``.T`` gives.
"""

from __future__ import annotations

from ..qdldl.types import CompiledQDLDL


def _t(x):
    if getattr(x, "ndim", 2) != 2:
        raise ValueError(f"expected a 2-D array, got shape {tuple(x.shape)}")
    return x.T


def values_for_compiled_qdldl(compiled: CompiledQDLDL, kkt_values):
    """(B, nnz_kkt) KKT values -> the layout ``compiled.factor`` takes."""
    if compiled.values_layout == "original_batch":
        return kkt_values
    if compiled.values_layout == "original_symbolic":
        return _t(kkt_values)
    raise ValueError(f"unknown QDLDL values layout {compiled.values_layout}")


def rhs_for_compiled_qdldl(compiled: CompiledQDLDL, rhs):
    """(B, n) right-hand sides -> the layout ``compiled.solve`` takes."""
    if compiled.rhs_layout == "batch":
        return rhs
    if compiled.rhs_layout == "symbolic":
        return _t(rhs)
    raise ValueError(f"unknown QDLDL rhs layout {compiled.rhs_layout}")


def batch_from_compiled_qdldl(compiled: CompiledQDLDL, values):
    """A solve result in ``compiled``'s layout -> batch-first (B, n)."""
    if compiled.rhs_layout == "symbolic":
        return _t(values)
    return values


__all__ = ["batch_from_compiled_qdldl", "rhs_for_compiled_qdldl", "values_for_compiled_qdldl"]
