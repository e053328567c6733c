"""Fixed-pattern SQP for batched nonlinear MPC on the HIP OSQP backend."""

from .line_search import (
    LINE_SEARCH_CONSTRAINT,
    LINE_SEARCH_COST,
    LINE_SEARCH_COST_OR_CONSTRAINT,
    LINE_SEARCH_MAX_ITER,
    FilterLineSearchResult,
    FilterLineSearchSettings,
    constraint_violation,
    filter_line_search_from_evaluations,
    make_step_lengths,
)
from .sparse_mpc import (
    CompiledSparseMPCSQP,
    SparseMPCPlan,
    SparseMPCProblem,
    SQPGraph,
    SQPLinearization,
    SQPLineSearchStepResult,
    SQPSolveResult,
    SQPStepResult,
    build_sparse_mpc_plan,
    compile_sparse_mpc_sqp,
)
from .stage import TorchStageFunction

__all__ = [
    "CompiledSparseMPCSQP",
    "FilterLineSearchResult",
    "FilterLineSearchSettings",
    "LINE_SEARCH_CONSTRAINT",
    "LINE_SEARCH_COST",
    "LINE_SEARCH_COST_OR_CONSTRAINT",
    "LINE_SEARCH_MAX_ITER",
    "SQPGraph",
    "SQPLineSearchStepResult",
    "SQPLinearization",
    "SQPSolveResult",
    "SQPStepResult",
    "SparseMPCPlan",
    "SparseMPCProblem",
    "TorchStageFunction",
    "build_sparse_mpc_plan",
    "compile_sparse_mpc_sqp",
    "constraint_violation",
    "filter_line_search_from_evaluations",
    "make_step_lengths",
]
