"""CPU oracle of the SQP step.  TEST INFRASTRUCTURE ONLY (tests and benchmark
checks; never called by the product path).

* Line search: numpy restatement of warpmpc/jax_sqp/line_search.py:45-169
  (step ladder, scaled bound violation, filter / Armijo / constraint
  acceptance, first accepted step or the last one with reason MAX_ITER).
* SQP step in float64: the NLP of a stage-ordered problem is evaluated by
  concatenating each stage's (cost, g, lower, upper) (sparse_mpc.py:840-854);
  its QP is assembled from DENSE derivatives of the whole NLP (total cost
  gradient and Hessian, constraint Jacobian, by torch.autograd.functional in
  float64) read at the plan's P / A patterns -- independent of the product's
  per-stage sparse derivatives and scatter positions -- and solved by the
  FP64 ADMM oracle from a warm start; then the line search
  (sparse_mpc.py:997-1053).
"""

from __future__ import annotations

import numpy as np

LINE_SEARCH_MAX_ITER = 0
LINE_SEARCH_CONSTRAINT = 1
LINE_SEARCH_COST = 2
LINE_SEARCH_COST_OR_CONSTRAINT = 3


def make_step_lengths(decay: float = 0.5, step_min: float = 0.1) -> np.ndarray:
    steps = []
    s = 1.0
    while s + max(1e-12, 1e-12 * step_min) >= step_min:
        steps.append(s)
        s *= decay
    return np.asarray(steps or [1.0])


def constraint_violation(lower, g, upper, scale: float) -> np.ndarray:
    lower, g, upper = (np.asarray(v, dtype=np.float64) for v in (lower, g, upper))
    return scale * np.sqrt(np.sum(np.maximum(lower - g, 0.0) ** 2, axis=-1)
                           + np.sum(np.minimum(upper - g, 0.0) ** 2, axis=-1))


def filter_line_search(*, step_lengths, baseline_cost, baseline_violation, armijo_metric, candidate_costs,
                       candidate_violations, g_max=1e-1, g_min=1e-3, gamma_c=1e-3, armijo_factor=1e-4,
                       uses_trial_cost=True):
    """Per instance (step, reason, accepted, index) from (n_steps, B) candidates."""
    cc = np.asarray(candidate_costs, dtype=np.float64)
    cv = np.asarray(candidate_violations, dtype=np.float64)
    k, b = cc.shape
    steps = np.asarray(step_lengths, dtype=np.float64)
    out_step = np.empty(b)
    out_reason = np.empty(b, dtype=np.int64)
    out_ok = np.empty(b, dtype=bool)
    out_idx = np.empty(b, dtype=np.int64)
    for j in range(b):
        c0, v0, dm = float(baseline_cost[j]), float(baseline_violation[j]), float(armijo_metric[j])
        chosen = None
        for i in range(k):
            c, v = cc[i, j], cv[i, j]
            if v > g_max:
                reason, ok = LINE_SEARCH_CONSTRAINT, v < (1.0 - gamma_c) * v0
            elif v < g_min and v0 < g_min and dm < 0.0:
                lhs = c if uses_trial_cost else c0
                reason, ok = LINE_SEARCH_COST, lhs < c0 + armijo_factor * dm
            else:
                reason, ok = LINE_SEARCH_COST_OR_CONSTRAINT, (c < c0 - gamma_c * v0) or (v < (1.0 - gamma_c) * v0)
            if ok:
                chosen = (i, reason)
                break
        if chosen is None:
            out_idx[j], out_reason[j], out_ok[j] = k - 1, LINE_SEARCH_MAX_ITER, False
        else:
            out_idx[j], out_reason[j], out_ok[j] = chosen[0], chosen[1], True
        out_step[j] = steps[out_idx[j]]
    return out_step, out_reason, out_ok, out_idx


class OracleSQP:
    """Float64 SQP step of a stage-ordered problem on the CPU."""

    def __init__(self, problem, plan):
        self.problem = problem
        self.plan = plan

    def _torch(self):
        import torch

        return torch

    def nlp(self, z, params):
        """(cost, g, lower, upper) of one instance as float64 torch tensors."""
        th = self._torch()
        pl = self.plan
        costs, gs, los, ups = [], [], [], []
        for i, st in enumerate(self.problem.stages):
            a = pl.assemblies[i]
            zs = z[a.z_offset:a.z_offset + st.z_dim]
            ps = params[a.param_offset:a.param_offset + st.param_dim]
            if st.has_next:
                c, g, lo, up = st.fn(zs, z[a.next_z_offset:a.next_z_offset + st.next_z_dim], ps)
            else:
                c, g, lo, up = st.fn(zs, ps)
            costs.append(c)
            gs.append(g)
            los.append(lo)
            ups.append(up)
        return th.stack(costs).sum(), th.cat(gs), th.cat(los), th.cat(ups)

    def linearize(self, z, params):
        """QP data (p_values, a_values, q, l, u) and (cost, g, lower, upper), float64 numpy."""
        th = self._torch()
        from torch.autograd.functional import hessian, jacobian

        zt = th.as_tensor(np.asarray(z, dtype=np.float64))
        pt = th.as_tensor(np.asarray(params, dtype=np.float64))
        cost, g, lo, up = self.nlp(zt, pt)
        grad = jacobian(lambda v: self.nlp(v, pt)[0], zt).numpy()
        hess = hessian(lambda v: self.nlp(v, pt)[0], zt).numpy()
        jac = jacobian(lambda v: self.nlp(v, pt)[1], zt).numpy()
        pp, ap = self.plan.p_pattern, self.plan.a_pattern
        pc = np.repeat(np.arange(pp.shape[1]), np.diff(pp.indptr))
        ac = np.repeat(np.arange(ap.shape[1]), np.diff(ap.indptr))
        p_values = hess[pp.indices, pc]
        a_values = jac[ap.indices, ac]
        g, lo, up = g.detach().numpy(), lo.detach().numpy(), up.detach().numpy()
        return (p_values, a_values, grad, lo - g, up - g), (float(cost), g, lo, up)

    def step(self, z, params, warm=None, *, beta=None, scale=0.1, decay=0.5, step_min=0.1,
             uses_trial_cost=True, max_iter=None):
        """One SQP step per instance (rows of z/params).  beta given: the fixed
        step z + beta d; else the line-search step.  Returns a dict of arrays."""
        from oracle import oracle

        z = np.atleast_2d(np.asarray(z, dtype=np.float64))
        params = np.atleast_2d(np.asarray(params, dtype=np.float64))
        b = z.shape[0]
        qp = [self.linearize(z[i], params[i]) for i in range(b)]
        pv, av, q, l, u = (np.stack([d[0][k] for d in qp]) for k in range(5))
        base = [d[1] for d in qp]
        orc = oracle.OracleOSQP(self.plan.osqp_plan, max_iter=max_iter)
        if warm is None:
            x, qz, qy, prim, dual, obj = orc.solve(pv, av, q, l, u)
        else:
            x, qz, qy, prim, dual, obj = orc.solve(pv, av, q, l, u, x0=warm[0], z0=warm[1], y0=warm[2])
        out = dict(direction=x, z=qz, y=qy, p_values=pv, a_values=av, q=q, l=l, u=u,
                   cost=np.array([c[0] for c in base]))
        if beta is not None:
            out["z_next"] = z + beta * x
            return out
        steps = make_step_lengths(decay, step_min)
        v0 = np.array([constraint_violation(c[2], c[1], c[3], scale) for c in base])
        metric = np.sum(q * x, axis=1)
        cc = np.empty((steps.size, b))
        cv = np.empty((steps.size, b))
        th = self._torch()
        for i, s in enumerate(steps):
            for j in range(b):
                c, g, lo, up = self.nlp(th.as_tensor(z[j] + s * x[j]), th.as_tensor(params[j]))
                cc[i, j] = float(c)
                cv[i, j] = constraint_violation(lo.numpy(), g.numpy(), up.numpy(), scale)
        step, reason, ok, idx = filter_line_search(
            step_lengths=steps, baseline_cost=out["cost"], baseline_violation=v0, armijo_metric=metric,
            candidate_costs=cc, candidate_violations=cv, uses_trial_cost=uses_trial_cost)
        out.update(z_next=z + step[:, None] * x, step_length=step, reason=reason, accepted=ok,
                   candidate_costs=cc, candidate_violations=cv, baseline_violation=v0, metric=metric)
        return out


__all__ = ["OracleSQP", "constraint_violation", "filter_line_search", "make_step_lengths"]
