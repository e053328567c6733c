"""Stage functions of a fixed-pattern nonlinear MPC, differentiated with torch.func.

``TorchStageFunction`` plays the role of the reference's CasADi-backed
``CasadiStageFunction`` (``warpmpc/jax_sqp/casadi_stage.py:41-191``): a stage
returns ``(cost, g, lower, upper)`` from ``(z_stage, z_next, p_stage)`` (or
``(z_stage, p_stage)`` for the terminal stage), and the SQP needs, with a
fixed local sparsity, the cost gradient, the upper triangle of the cost
Hessian and the Jacobian of ``g``, all with respect to the stage's local
variables ``[z_stage, z_next]``.

CasADi is not part of this build.  Stage functions are torch functions that
accept any leading batch shape (the last axis holds the stage's variables);
derivatives come from ``torch.func`` (vmap of jacrev / hessian), so they run
on the GPU and stay differentiable.  The fixed sparsity is the union of the
nonzero derivative entries at several random float64 points: the structural
pattern of any generic stage, found the way CasADi's symbolic sparsity would
find it.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable

import numpy as np


def _torch():
    import torch

    return torch


_CONSTS: dict = {}


def device_const(arr: np.ndarray, dtype, device):
    """A device copy of a constant host array, made once per (array, dtype,
    device): steps issue no host-to-device copies after their first call, so
    they can be captured in a HIP graph."""
    key = (id(arr), str(dtype), str(device))
    hit = _CONSTS.get(key)
    if hit is None or hit[0] is not arr:
        hit = (arr, _torch().as_tensor(arr, dtype=dtype, device=device))
        _CONSTS[key] = hit
    return hit[1]


@dataclass(frozen=True)
class TorchStageFunction:
    name: str
    fn: Callable
    has_next: bool
    z_dim: int
    next_z_dim: int
    param_dim: int
    constraint_dim: int
    grad_cols: np.ndarray
    hess_rows: np.ndarray
    hess_cols: np.ndarray
    jac_rows: np.ndarray
    jac_cols: np.ndarray
    # cost-Hessian entries when equal at every probe (a quadratic cost): then
    # they are used as constants and the Hessian is not differentiated out
    hess_const: np.ndarray | None = None

    @classmethod
    def from_function(cls, fn: Callable, *, z_dim: int, param_dim: int, next_z_dim: int | None = None,
                      name: str = "stage", probes: int = 4, seed: int = 0) -> "TorchStageFunction":
        """``next_z_dim=None`` makes a terminal stage ``fn(z, p)``; otherwise ``fn(z, zn, p)``."""
        th = _torch()
        has_next = next_z_dim is not None
        zn_dim = int(next_z_dim) if has_next else 0
        local = int(z_dim) + zn_dim
        rng = np.random.default_rng(seed)
        out = _call(fn, has_next, int(z_dim), th.zeros(local, dtype=th.float64), th.zeros(param_dim, dtype=th.float64))
        if out[0].dim() != 0:
            raise ValueError("stage cost must be a scalar per instance")
        c = int(out[1].numel())
        if out[1].dim() != 1 or out[2].shape != out[1].shape or out[3].shape != out[1].shape:
            raise ValueError("stage g, lower and upper must be vectors of one length")
        grad_nz = np.zeros(local, dtype=bool)
        hess_nz = np.zeros((local, local), dtype=bool)
        jac_nz = np.zeros((c, local), dtype=bool)
        hess_seen = []
        for _ in range(max(2, probes)):
            loc = th.as_tensor(rng.standard_normal(local))
            p = th.as_tensor(rng.standard_normal(param_dim))
            grad, hess, jac = _derivatives(fn, has_next, int(z_dim), loc[None], p[None])
            grad_nz |= grad[0].numpy() != 0
            hess_nz |= hess[0].numpy() != 0
            jac_nz |= jac[0].numpy() != 0
            hess_seen.append(hess[0].numpy())
        hess_nz |= hess_nz.T
        hr, hc = np.nonzero(np.triu(hess_nz))
        order = np.lexsort((hr, hc))  # column-major, like a CSC triplet walk
        h_entries = np.stack([h[hr[order], hc[order]] for h in hess_seen])
        hess_const = h_entries[0].copy() if np.all(h_entries == h_entries[0]) else None
        jr, jcol = np.nonzero(jac_nz)
        jorder = np.lexsort((jr, jcol))
        return cls(
            name=name,
            fn=fn,
            has_next=has_next,
            z_dim=int(z_dim),
            next_z_dim=zn_dim,
            param_dim=int(param_dim),
            constraint_dim=c,
            grad_cols=np.flatnonzero(grad_nz).astype(np.int32),
            hess_rows=hr[order].astype(np.int32),
            hess_cols=hc[order].astype(np.int32),
            jac_rows=jr[jorder].astype(np.int32),
            jac_cols=jcol[jorder].astype(np.int32),
            hess_const=hess_const,
        )

    def values(self, z, zn, p):
        """(cost, g, lower, upper) for batched inputs (zn ignored for a terminal stage)."""
        th = _torch()
        loc = th.cat([z, zn], dim=-1) if self.has_next else z
        return _call(self.fn, self.has_next, self.z_dim, loc, p)

    def sparse_values(self, z, zn, p):
        """(cost, g, lower, upper, grad entries, hess entries, jac entries); inputs
        (..., dim) with any leading shape, outputs (..., entries)."""
        th = _torch()
        loc = th.cat([z, zn], dim=-1) if self.has_next else z
        lead = loc.shape[:-1]
        loc2 = loc.reshape(-1, loc.shape[-1])
        p2 = p.reshape(-1, p.shape[-1])
        cost, g, lower, upper = _call(self.fn, self.has_next, self.z_dim, loc2, p2)
        constant = self.hess_const is not None
        grad, hess, jac = _derivatives(self.fn, self.has_next, self.z_dim, loc2, p2, hessian_too=not constant)
        dev = loc.device
        idx = lambda a: device_const(a, th.long, dev)  # noqa: E731
        ge = grad[:, idx(self.grad_cols)]
        if constant:
            he = device_const(self.hess_const, loc.dtype, dev).expand(loc2.shape[0], -1)
        else:
            he = hess[:, idx(self.hess_rows), idx(self.hess_cols)]
        je = jac[:, idx(self.jac_rows), idx(self.jac_cols)]
        shape = lambda t: t.reshape(*lead, *t.shape[1:])  # noqa: E731
        return tuple(shape(t) for t in (cost, g, lower, upper, ge, he, je))


def _call(fn, has_next: bool, z_dim: int, loc, p):
    if has_next:
        return fn(loc[..., :z_dim], loc[..., z_dim:], p)
    return fn(loc, p)


def _derivatives(fn, has_next: bool, z_dim: int, loc, p, hessian_too: bool = True):
    """Dense cost gradient (B, L), cost Hessian (B, L, L) or None, and g
    Jacobian (B, c, L)."""
    from torch.func import hessian, jacrev, vmap

    def cost1(x, q):
        return _call(fn, has_next, z_dim, x, q)[0]

    def g1(x, q):
        return _call(fn, has_next, z_dim, x, q)[1]

    grad = vmap(jacrev(cost1))(loc, p)
    hess = vmap(hessian(cost1))(loc, p) if hessian_too else None
    jac = vmap(jacrev(g1))(loc, p)
    return grad, hess, jac


__all__ = ["TorchStageFunction", "device_const"]
