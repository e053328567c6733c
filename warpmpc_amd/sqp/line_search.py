"""Fixed-candidate filter line search of the SQP step (batched, torch).

Restates ``warpmpc/jax_sqp/line_search.py:18-169``: a static ladder of step
lengths (1, decay, decay², ... down to ``line_search_step_min``); for every
candidate the cost and the scaled constraint violation; per instance the
first candidate accepted by the reason that applies to it:

* violation above ``g_max``: accept only a violation decrease by (1 - gamma_c);
* violation below ``g_min`` at both points and a descent direction: Armijo
  on the cost (trial cost, or the baseline cost when
  ``line_search_cost_accept_uses_trial_cost`` is False);
* otherwise the filter: lower cost by gamma_c times the violation, or a lower
  violation;

and the last (shortest) step with reason MAX_ITER when none is accepted.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import NamedTuple

import numpy as np

LINE_SEARCH_MAX_ITER = 0
LINE_SEARCH_CONSTRAINT = 1
LINE_SEARCH_COST = 2
LINE_SEARCH_COST_OR_CONSTRAINT = 3


@dataclass(frozen=True)
class FilterLineSearchSettings:
    """Parameters of the fixed-candidate filter line search (same defaults)."""

    line_search_g_max: float = 1e-1
    line_search_g_min: float = 1e-3
    line_search_gamma_c: float = 1e-3
    line_search_armijo_factor: float = 1e-4
    line_search_step_decay: float = 0.5
    line_search_step_min: float = 0.1
    line_search_constraint_scale: float = 0.1
    line_search_cost_accept_uses_trial_cost: bool = True


class FilterLineSearchResult(NamedTuple):
    step_length: object
    cost: object
    constraint_violation: object
    reason: object
    accepted: object
    candidate_costs: object
    candidate_constraint_violations: object
    candidate_acceptance: object


def make_step_lengths(settings: FilterLineSearchSettings) -> np.ndarray:
    """The step ladder 1, decay, decay², ... while >= step_min (at least [1])."""
    if not 0.0 < settings.line_search_step_decay < 1.0:
        raise ValueError("line_search_step_decay must be in (0, 1)")
    if settings.line_search_step_min <= 0.0:
        raise ValueError("line_search_step_min must be positive")
    out = []
    step = 1.0
    slack = max(1e-12, 1e-12 * settings.line_search_step_min)
    while step + slack >= settings.line_search_step_min:
        out.append(step)
        step *= settings.line_search_step_decay
    return np.asarray(out or [1.0], dtype=np.float64)


def _torch():
    import torch

    return torch


def constraint_violation(lower, g, upper, scale: float):
    """scale * sqrt(||max(l - g, 0)||² + ||min(u - g, 0)||²) per row of (B, m) arrays."""
    th = _torch()
    lower, g, upper = (th.as_tensor(v) for v in (lower, g, upper))
    below = th.clamp_min(lower - g, 0.0)
    above = th.clamp_max(upper - g, 0.0)
    return scale * th.sqrt((below * below).sum(dim=-1) + (above * above).sum(dim=-1))


def filter_line_search_from_evaluations(*, settings: FilterLineSearchSettings, step_lengths, baseline_cost,
                                        baseline_constraint_violation, armijo_descent_metric, candidate_costs,
                                        candidate_constraint_violations) -> FilterLineSearchResult:
    """First accepted step per instance; candidate arrays are (n_steps, B)."""
    th = _torch()
    cc = th.as_tensor(candidate_costs)
    dt, dev = cc.dtype, cc.device
    as_t = lambda v: th.as_tensor(v, dtype=dt, device=dev)  # noqa: E731
    steps = as_t(step_lengths)
    c0 = as_t(baseline_cost)[None, :]
    v0 = as_t(baseline_constraint_violation)[None, :]
    dmetric = as_t(armijo_descent_metric)[None, :]
    cv = as_t(candidate_constraint_violations)
    s = settings
    keep_violation = cv < (1.0 - s.line_search_gamma_c) * v0
    by_constraint = cv > s.line_search_g_max
    by_cost = (cv < s.line_search_g_min) & (v0 < s.line_search_g_min) & (dmetric < 0.0)
    lhs = cc if s.line_search_cost_accept_uses_trial_cost else c0.expand_as(cc)
    armijo = lhs < c0 + s.line_search_armijo_factor * dmetric
    filt = (cc < c0 - s.line_search_gamma_c * v0) | keep_violation
    accepted = th.where(by_constraint, keep_violation, th.where(by_cost, armijo, filt))
    reason = th.where(by_constraint, th.full_like(cc, LINE_SEARCH_CONSTRAINT, dtype=th.int64),
                      th.where(by_cost, th.full_like(cc, LINE_SEARCH_COST, dtype=th.int64),
                               th.full_like(cc, LINE_SEARCH_COST_OR_CONSTRAINT, dtype=th.int64)))
    any_ok = accepted.any(dim=0)
    first = th.argmax(accepted.to(th.int32), dim=0)
    pick = th.where(any_ok, first, th.full_like(first, steps.shape[0] - 1))
    cols = th.arange(cc.shape[1], device=dev)
    return FilterLineSearchResult(
        step_length=steps[pick],
        cost=cc[pick, cols],
        constraint_violation=cv[pick, cols],
        reason=th.where(any_ok, reason[pick, cols], th.full_like(pick, LINE_SEARCH_MAX_ITER)),
        accepted=any_ok,
        candidate_costs=cc,
        candidate_constraint_violations=cv,
        candidate_acceptance=accepted,
    )


__all__ = [
    "FilterLineSearchResult",
    "FilterLineSearchSettings",
    "LINE_SEARCH_CONSTRAINT",
    "LINE_SEARCH_COST",
    "LINE_SEARCH_COST_OR_CONSTRAINT",
    "LINE_SEARCH_MAX_ITER",
    "constraint_violation",
    "filter_line_search_from_evaluations",
    "make_step_lengths",
]
