"""The error measure of the GPU parity tests against the FP64 oracle.

Per instance (row), the norm-wise relative error ``|v - ref|_inf / |ref|_inf``.
A row whose reference is all but zero (a dual with every constraint inactive,
the gradient of an inactive bound) is measured against ``floor_frac`` times the
median ``|ref|_inf`` of the batch: a floor set by the data's own scale, not an
absolute 1, so small outputs are still held to a relative bound.
"""

from __future__ import annotations

import numpy as np


def rel_err(v, ref, floor_frac: float = 1e-3, row_floor=None) -> float:
    """``row_floor`` (one value per row) replaces the batch-median floor."""
    ref = np.asarray(ref, dtype=np.float64)
    ref = ref.reshape(ref.shape[0], -1) if ref.ndim else ref.reshape(1, 1)
    v = np.asarray(v, dtype=np.float64).reshape(ref.shape)
    if ref.size == 0:
        return 0.0
    scale = np.max(np.abs(ref), axis=1)
    tiny = float(np.finfo(np.float32).tiny)
    if row_floor is None:
        floor = floor_frac * max(float(np.median(scale)), tiny)
    else:
        floor = np.maximum(np.asarray(row_floor, dtype=np.float64).reshape(-1), tiny)
    return float(np.max(np.max(np.abs(v - ref), axis=1) / np.maximum(scale, floor)))


def _rows(v):
    v = np.asarray(v, dtype=np.float64)
    return v[:, None] if v.ndim == 1 else v


def osqp_errors(out, ref, q=None) -> dict:
    """Errors of an OSQP result ``(x, z, y[, prim_res, dual_res, obj_val])``
    (numpy, one row per instance) against the FP64 oracle's.

    All are measured per instance against its iterate scale ``S =
    max(|x64|_inf, |z64|_inf, |y64|_inf)``, the magnitude that float32
    rounding in the ADMM updates is relative to (about 1e-7 of S per
    operation, accumulating over the iterations):

    * x, z, y norm-wise relative, ``|v - v64|_inf / max(|v64|_inf, 1e-2 S)``:
      a vector below 1 % of its instance's scale (the multipliers of a QP
      with every constraint all but inactive) is held to 1e-6 S absolutely at
      the 1e-4 tolerance, about ten float32 ulps of S;
    * the residuals, differences of the iterates, relative to S;
    * the objective ``0.5 x'Px + q'x``, which cancels when its two terms nearly
      balance, relative to the sum of their magnitudes, ``|q'x64| + |obj64 -
      q'x64|`` (``q``, one row per instance, is then required).
    """
    names = ("x", "z", "y", "prim_res", "dual_res", "obj_val")
    iter_scale = np.max([np.max(np.abs(_rows(r)), axis=1) for r in ref[:3]], axis=0)
    errs = {}
    for name, o, r in zip(names[:3], out[:3], ref[:3]):
        errs[name] = rel_err(o, r, row_floor=1e-2 * iter_scale)
    if len(out) > 3:
        if q is None:
            raise ValueError("the objective's error measure needs q")
        tiny = np.finfo(np.float32).tiny
        for name, o, r in zip(names[3:5], out[3:5], ref[3:5]):
            d = np.abs(np.asarray(o, dtype=np.float64) - np.asarray(r, dtype=np.float64))
            errs[name] = float(np.max(d / np.maximum(iter_scale, tiny)))
        x64 = _rows(ref[0])
        obj64 = np.asarray(ref[5], dtype=np.float64).reshape(-1)
        qx = np.sum(np.broadcast_to(np.asarray(q, dtype=np.float64), x64.shape) * x64, axis=1)
        term_scale = np.abs(qx) + np.abs(obj64 - qx)
        d = np.abs(np.asarray(out[5], dtype=np.float64).reshape(-1) - obj64)
        errs["obj_val"] = float(np.max(d / np.maximum(term_scale, tiny)))
    return errs


def adjoint_errors(got, ref, names) -> dict:
    """Errors of adjoint outputs (dq, dl, du, dP, dA, ... : numpy, one row per
    instance) against the FP64 oracle's, each norm-wise relative per instance
    and floored at 1 % of the instance's largest output ``max_k |ref_k|_inf``:
    an output that is all but zero for an instance (every derivative of a
    decision fixed by an active bound, except that bound's own) is held to
    1e-6 of that instance's gradient scale at the 1e-4 tolerance."""
    rows = [np.asarray(r, dtype=np.float64).reshape(np.shape(r)[0], -1) for r in ref]
    scale = np.max([np.max(np.abs(r), axis=1) if r.shape[1] else np.zeros(r.shape[0]) for r in rows], axis=0)
    return {name: rel_err(g, r, row_floor=1e-2 * scale) for name, g, r in zip(names, got, rows)}
