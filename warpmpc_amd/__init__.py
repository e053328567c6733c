"""MI355X-native batched fixed-pattern QP core: QDLDL LDLᵀ, OSQP ADMM and its adjoint."""

__version__ = "0.1.0"
