"""OSQP adjoint derivative oracle (CPU only).

The oracle restates warpmpc/jax_osqp/solver.py:1035-1192.  That formula is the
vector-Jacobian product of the map (P, A, q, l, u) -> (x, |y|):

* ``dx`` is the cotangent of the primal solution;
* ``dy`` is the cotangent of the split multipliers (y_l, y_u) = (max(-y, 0),
  max(y, 0)), i.e. of |y| -- rows with y < 0 (lower-active inequalities and
  equalities with a negative multiplier) take the cotangent with the sign of
  -y.  ``solver.py:1124-1128`` (rhs ``-dy`` on lower rows, ``d_nu`` flipped when
  y < 0) is where this shows;
* ``dP`` is the derivative with respect to the symmetric matrix, returned on
  the upper-triangle entries with OSQP's convention: diagonal entries as is,
  off-diagonal entries at one half of the derivative with respect to the
  stored value (which moves both P_ij and P_ji).

Pins:
* FP64 central finite differences (h = 1e-6) of xb.x + yb.|y| on tightly
  converged FP64 ADMM solves, over every output, on QPs whose active sets
  cover upper-active, lower-active, two-sided and equality rows of both
  multiplier signs;
* an independent dense active-set implicit derivative (numpy.linalg.solve on
  the KKT system of the active constraints; no adjoint matrix, no QDLDL
  tables), which has no finite-difference noise;
* the iterative refinement: with refine = 0 the result is the perturbed
  (eps = 1e-6) solve and visibly differs; with the default 100 steps the
  adjoint solution solves the unperturbed system;
* shared-operand (stride 0) batches equal per-instance batches.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import oracle
from warpmpc_amd.osqp import OSQPSettings, build_osqp_plan

INF = 1e30
TIGHT = OSQPSettings(max_iter=20000, rho=0.1, sigma=1e-6, alpha=1.6, eps_abs=1e-13, eps_rel=1e-13)


def _plan(p, a, l, u):
    return build_osqp_plan(sp.csc_matrix(p), sp.csc_matrix(a), l, u, TIGHT, derivatives=True)


def _solve(orc, plan, pv, av, q, l, u):
    x, _, y, prim, dual, _ = orc.solve(pv, av, q, l, u)
    return x, y, prim, dual


def _fd_gradients(plan, q, l, u, xb, yb, h=1e-6):
    """Central differences of xb.x + yb.|y| w.r.t. every problem value, one batch."""
    orc = oracle.OracleOSQP(plan)
    pv0, av0 = plan.p_upper.data.copy(), plan.a_matrix.data.copy()
    n, m, npv, nav = plan.n, plan.m, pv0.size, av0.size
    cases = []  # (field, index, sign)
    for field, size in (("p", npv), ("a", nav), ("q", n), ("l", m), ("u", m)):
        for i in range(size):
            cases += [(field, i, 1.0), (field, i, -1.0)]
    bsz = len(cases)
    pv = np.tile(pv0, (bsz, 1))
    av = np.tile(av0, (bsz, 1))
    qq = np.tile(q, (bsz, 1))
    ll = np.tile(l, (bsz, 1))
    uu = np.tile(u, (bsz, 1))
    target = {"p": pv, "a": av, "q": qq, "l": ll, "u": uu}
    for k, (field, i, s) in enumerate(cases):
        target[field][k, i] += s * h
        if field in ("l", "u") and l[i] == u[i]:
            # an equality row moves both bounds together: its value is l = u
            other = uu if field == "l" else ll
            other[k, i] += s * h
    x, y, *_ = _solve(orc, plan, pv, av, qq, ll, uu)
    loss = x @ xb + np.abs(y) @ yb
    grads = {}
    k = 0
    for field, size in (("p", npv), ("a", nav), ("q", n), ("l", m), ("u", m)):
        g = np.zeros(size)
        for i in range(size):
            g[i] = (loss[k] - loss[k + 1]) / (2 * h)
            k += 2
        grads[field] = g
    return grads


def _dense_active_set(plan, q, l, u, x, y, xb, yb, tol=1e-7):
    """Implicit derivative of (x, |y|) from the KKT system of the active rows."""
    n, m = plan.n, plan.m
    pu = plan.p_upper.toarray()
    p_full = np.triu(pu) + np.triu(pu, 1).T
    a = plan.a_matrix.toarray()
    act = np.flatnonzero(np.abs(y) > tol)
    k = np.block([[p_full, a[act].T], [a[act], np.zeros((act.size, act.size))]])
    s = np.sign(y[act])
    w = np.linalg.solve(k.T, np.concatenate([xb, s * yb[act]]))
    wx, wy = w[:n], w[n:]
    full_y = np.zeros(m)
    full_y[act] = wy
    dq = -wx
    # b of an active row is u (y > 0) or l (y < 0); an equality row splits it
    dl = np.where(y < 0, full_y, 0.0)
    du = np.where(y > 0, full_y, 0.0)
    pr, pc = plan.p_upper.nonzero()
    order = np.lexsort((pr, pc))  # CSC entry order
    pr, pc = pr[order], pc[order]
    dp = -0.5 * (wx[pr] * x[pc] + wx[pc] * x[pr])
    ar, ac = plan.a_matrix.nonzero()
    order = np.lexsort((ar, ac))
    ar, ac = ar[order], ac[order]
    # [x; y_act] solves K [x; y_act] = [-q; b]; K's A block appears twice
    da = -(full_y[ar] * x[ac] + wx[ac] * y[ar])
    return {"q": dq, "l": dl, "u": du, "p": dp, "a": da}


def _adjoint(plan, x, y, xb, yb, l, u, refine=100):
    adj = oracle.OracleAdjoint(plan, refine=refine)
    sol, ryl, ryu, dq, dl, du, dp, da = adj.compute(plan.p_upper.data[None], plan.a_matrix.data[None], l, u,
                                                    x, y, xb, yb)
    return {"sol": sol[0], "q": dq[0], "l": dl[0], "u": du[0], "p": dp[0], "a": da[0], "ryl": ryl[0],
            "ryu": ryu[0]}


P3 = np.array([[4.0, 1.0, 0.0], [1.0, 2.0, 0.5], [0.0, 0.5, 3.0]])
A_EQ = np.array([[1.0, 1.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
CASES = {
    # equality with y < 0, two-sided row active at its upper bound
    "eq_neg_upper": (P3, A_EQ, [1.0, -0.2, -INF, 0.0], [1.0, 0.3, 0.1, INF], [-1.0, 2.0, -0.5]),
    # equality with y > 0
    "eq_pos": (P3, A_EQ, [-1.0, -0.2, -INF, 0.0], [-1.0, 0.3, 0.1, INF], [-1.0, 2.0, -0.5]),
    # inequalities only, upper- and lower-active rows
    "ineq_mixed": (P3, np.eye(3), [-0.2, -INF, 0.5], [0.3, 0.1, INF], [-1.0, 2.0, -0.5]),
    # every row active, lower-active rows on two-sided and one-sided bounds
    "ineq_lower": (P3, np.eye(3), [-0.2, -INF, 0.5], [0.3, 0.1, INF], [2.0, -3.0, 1.5]),
}


def _random_qp(seed):
    rng = np.random.default_rng(seed)
    n, m = 6, 9
    g = rng.standard_normal((n, n))
    p = g @ g.T + n * np.eye(n)
    p[np.abs(p) < 0.8] = 0.0
    p = p + np.eye(n)
    a = rng.standard_normal((m, n))
    a[rng.random((m, n)) < 0.4] = 0.0
    a[np.arange(m), rng.integers(0, n, m)] = 1.0
    q = rng.standard_normal(n) * 3
    l = -rng.random(m) * 0.5
    u = rng.random(m) * 0.5
    l[0] = u[0] = 0.2          # equality
    l[1] = -INF                # upper only
    u[2] = INF                 # lower only
    return p, a, l, u, q


def _check_strict(x, y, plan, l, u, margin=1e-4):
    """Strict complementarity: the active set is stable under the FD step."""
    ax = plan.a_matrix.toarray() @ x
    for i in range(plan.m):
        if l[i] == u[i]:
            continue
        slack = min(ax[i] - l[i], u[i] - ax[i])
        active = abs(y[i]) > margin and slack < 1e-8
        inactive = abs(y[i]) < 1e-8 and slack > margin
        assert active or inactive, (i, y[i], slack)


def _problem(name):
    if name.startswith("random"):
        return _random_qp(int(name.split("_")[1]))
    p, a, l, u, q = CASES[name]
    return np.asarray(p), np.asarray(a), np.asarray(l, float), np.asarray(u, float), np.asarray(q, float)


@pytest.mark.parametrize("name", list(CASES) + ["random_3", "random_11"])
def test_adjoint_matches_finite_differences(name):
    p, a, l, u, q = _problem(name)
    plan = _plan(np.triu(p), a, l, u)
    orc = oracle.OracleOSQP(plan)
    x, y, prim, dual = _solve(orc, plan, plan.p_upper.data[None], plan.a_matrix.data[None], q[None], l[None],
                              u[None])
    assert prim[0] < 1e-10 and dual[0] < 1e-10
    x, y = x[0], y[0]
    _check_strict(x, y, plan, l, u)
    rng = np.random.default_rng(7)
    xb, yb = rng.standard_normal(plan.n), rng.standard_normal(plan.m)
    got = _adjoint(plan, x, y, xb, yb, l, u)
    fd = _fd_gradients(plan, q, l, u, xb, yb)
    np.testing.assert_allclose(got["q"], fd["q"], atol=1e-7)
    two_sided = l < u
    np.testing.assert_allclose(np.where(two_sided, got["l"], 0), np.where(two_sided, fd["l"], 0), atol=1e-7)
    np.testing.assert_allclose(np.where(two_sided, got["u"], 0), np.where(two_sided, fd["u"], 0), atol=1e-7)
    # equality rows: the bound value is l = u, its derivative is dl + du
    eq = ~two_sided
    np.testing.assert_allclose((got["l"] + got["u"])[eq], fd["l"][eq], atol=1e-7)
    pr, pc = plan.derivative_plan.p_upper_rows, plan.derivative_plan.p_upper_cols
    np.testing.assert_allclose(got["p"] * np.where(pr == pc, 1.0, 2.0), fd["p"], atol=1e-7)
    np.testing.assert_allclose(got["a"], fd["a"], atol=1e-7)


@pytest.mark.parametrize("name", list(CASES) + ["random_3", "random_11", "random_19"])
def test_adjoint_matches_dense_active_set(name):
    p, a, l, u, q = _problem(name)
    plan = _plan(np.triu(p), a, l, u)
    orc = oracle.OracleOSQP(plan)
    x, y, *_ = _solve(orc, plan, plan.p_upper.data[None], plan.a_matrix.data[None], q[None], l[None], u[None])
    x, y = x[0], y[0]
    rng = np.random.default_rng(1)
    xb, yb = rng.standard_normal(plan.n), rng.standard_normal(plan.m)
    got = _adjoint(plan, x, y, xb, yb, l, u)
    ref = _dense_active_set(plan, q, l, u, x, y, xb, yb)
    for key in ("q", "l", "u", "p", "a"):
        np.testing.assert_allclose(got[key], ref[key], atol=1e-9, err_msg=key)


def test_refinement_removes_the_perturbation():
    p, a, l, u, q = _problem("random_3")
    plan = _plan(np.triu(p), a, l, u)
    orc = oracle.OracleOSQP(plan)
    x, y, *_ = _solve(orc, plan, plan.p_upper.data[None], plan.a_matrix.data[None], q[None], l[None], u[None])
    x, y = x[0], y[0]
    xb, yb = np.linspace(-1, 1, plan.n), np.linspace(1, -1, plan.m)
    ref = _dense_active_set(plan, q, l, u, x, y, xb, yb)
    raw = _adjoint(plan, x, y, xb, yb, l, u, refine=0)
    one = _adjoint(plan, x, y, xb, yb, l, u, refine=1)
    full = _adjoint(plan, x, y, xb, yb, l, u, refine=100)
    e_raw = np.abs(raw["q"] - ref["q"]).max()
    e_one = np.abs(one["q"] - ref["q"]).max()
    e_full = np.abs(full["q"] - ref["q"]).max()
    assert e_raw > 1e-8          # the eps = 1e-6 perturbation is visible
    assert e_one < e_raw * 1e-2  # each step contracts by ~eps
    assert e_full < 1e-11


def test_shared_and_per_instance_batches_agree():
    p, a, l, u, q = _problem("random_11")
    plan = _plan(np.triu(p), a, l, u)
    orc = oracle.OracleOSQP(plan)
    rng = np.random.default_rng(5)
    b = 6
    qb = q[None] + 0.3 * rng.standard_normal((b, plan.n))
    x, _, y, *_ = orc.solve(plan.p_upper.data[None], plan.a_matrix.data[None], qb, l[None], u[None])
    xb, yb = rng.standard_normal((b, plan.n)), rng.standard_normal((b, plan.m))
    adj = oracle.OracleAdjoint(plan)
    shared = adj.compute(plan.p_upper.data[None], plan.a_matrix.data[None], l, u, x, y, xb, yb)
    per = adj.compute(np.tile(plan.p_upper.data, (b, 1)), np.tile(plan.a_matrix.data, (b, 1)),
                      np.tile(l, (b, 1)), np.tile(u, (b, 1)), x, y, xb, yb)
    for s_, p_ in zip(shared, per):
        np.testing.assert_array_equal(s_, p_)
    for i in range(b):
        one = adj.compute(plan.p_upper.data[None], plan.a_matrix.data[None], l, u, x[i], y[i], xb[i], yb[i])
        for s_, o_ in zip(shared, one):
            np.testing.assert_array_equal(s_[i], o_[0])


def test_f32_adjoint_tracks_f64():
    p, a, l, u, q = _problem("random_3")
    plan = _plan(np.triu(p), a, l, u)
    orc = oracle.OracleOSQP(plan)
    x, _, y, *_ = orc.solve(plan.p_upper.data[None], plan.a_matrix.data[None], q[None], l[None], u[None])
    xb, yb = np.linspace(-1, 1, plan.n), np.linspace(1, -1, plan.m)
    adj = oracle.OracleAdjoint(plan, refine=10)
    r64 = adj.compute(plan.p_upper.data[None], plan.a_matrix.data[None], l, u, x, y, xb, yb)
    r32 = adj.compute(plan.p_upper.data[None], plan.a_matrix.data[None], l, u, x, y, xb, yb, dtype=np.float32)
    for a64, a32 in zip(r64[3:], r32[3:]):
        np.testing.assert_allclose(a32, a64, atol=2e-4 * max(1.0, np.abs(a64).max()))


def test_no_derivative_plan_is_an_error():
    p, a, l, u, q = _problem("eq_pos")
    plan = build_osqp_plan(sp.csc_matrix(np.triu(p)), sp.csc_matrix(a), l, u, TIGHT)
    with pytest.raises(ValueError):
        oracle.OracleAdjoint(plan)


# --------------------------------------------------------------------------- converged limit


@pytest.mark.parametrize("name", list(CASES) + ["random_3", "random_11", "random_19"])
def test_exact_adjoint_is_the_limit_of_the_refinement(name):
    """adjoint_exact (dense U solve) equals the C restatement with the
    refinement run to convergence, and the dense active-set derivative."""
    p, a, l, u, q = _problem(name)
    plan = _plan(np.triu(p), a, l, u)
    orc = oracle.OracleOSQP(plan)
    x, y, *_ = _solve(orc, plan, plan.p_upper.data[None], plan.a_matrix.data[None], q[None], l[None], u[None])
    x, y = x[0], y[0]
    rng = np.random.default_rng(2)
    xb, yb = rng.standard_normal(plan.n), rng.standard_normal(plan.m)
    ex = oracle.adjoint_exact(plan, plan.p_upper.data, plan.a_matrix.data, l, u, x, y, xb, yb)
    ref = _adjoint(plan, x, y, xb, yb, l, u, refine=400)
    for key, v in zip(("sol", "ryl", "ryu", "q", "l", "u", "p", "a"), ex):
        np.testing.assert_allclose(v[0], ref[key], atol=1e-9, rtol=1e-9, err_msg=key)
    dense = _dense_active_set(plan, q, l, u, x, y, xb, yb)
    for key, v in zip(("q", "l", "u", "p", "a"), ex[3:]):
        np.testing.assert_allclose(v[0], dense[key], atol=1e-9, err_msg=key)


def test_exact_adjoint_differs_from_a_short_refinement_on_an_ill_conditioned_qp():
    """With cond(U) large the reference's 100 refinement steps stop short of
    the limit; adjoint_exact is the limit (pinned above), so the two differ."""
    n = 4
    p = np.diag([1.0, 1e-3, 1.0, 1.0])
    a = np.array([[1.0, 1.0, 0.0, 0.0], [1.0, 1.0 + 1e-4, 0.0, 0.0], [0.0, 0.0, 1.0, 0.0]])
    l = np.array([1.0, -INF, -1.0])
    u = np.array([1.0, 1.0, 1.0])
    q = np.array([1.0, -2.0, 0.5, 0.3])
    plan = _plan(np.triu(p), a, l, u)
    orc = oracle.OracleOSQP(plan)
    x, y, *_ = _solve(orc, plan, plan.p_upper.data[None], plan.a_matrix.data[None], q[None], l[None], u[None])
    xb, yb = np.ones(n), np.ones(3)
    ex = oracle.adjoint_exact(plan, plan.p_upper.data, plan.a_matrix.data, l, u, x[0], y[0], xb, yb)
    r100 = _adjoint(plan, x[0], y[0], xb, yb, l, u, refine=100)
    r_long = _adjoint(plan, x[0], y[0], xb, yb, l, u, refine=20000)
    np.testing.assert_allclose(ex[3][0], r_long["q"], rtol=1e-6, atol=1e-9)
    assert np.abs(ex[3][0] - r100["q"]).max() > 1e3 * np.abs(ex[3][0] - r_long["q"]).max()


@pytest.mark.parametrize("seed", [3, 11, 19])
def test_clamped_slacks_give_the_active_set_derivative_at_an_infeasible_iterate(seed):
    """After 3 ADMM iterations rows with a positive multiplier violate their
    bounds.  adjoint_exact(clamp_slacks=True) there equals the dense
    active-set implicit derivative (active rows: y != 0), while the unclamped
    limit differs; at a converged point clamping changes nothing."""
    p, a, l, u, q = _problem(f"random_{seed}")
    early = build_osqp_plan(sp.csc_matrix(np.triu(p)), sp.csc_matrix(a), l, u,
                            OSQPSettings(max_iter=3, rho=0.1, sigma=1e-6, alpha=1.6), derivatives=True)
    x, _, y, *_ = oracle.OracleOSQP(early).solve(early.p_upper.data[None], early.a_matrix.data[None], q[None],
                                                 l[None], u[None])
    x, y = x[0], y[0]
    ax = a @ x
    assert np.any((l < u) & (((y > 0) & (ax > u)) | ((y < 0) & (ax < l))))
    rng = np.random.default_rng(seed)
    xb, yb = rng.standard_normal(early.n), rng.standard_normal(early.m)
    pv, av = early.p_upper.data, early.a_matrix.data
    clamped = oracle.adjoint_exact(early, pv, av, l, u, x, y, xb, yb, clamp_slacks=True)
    raw = oracle.adjoint_exact(early, pv, av, l, u, x, y, xb, yb)
    dense = _dense_active_set(early, q, l, u, x, y, xb, yb, tol=0.0)
    for key, v in zip(("q", "l", "u", "p", "a"), clamped[3:]):
        np.testing.assert_allclose(v[0], dense[key], atol=1e-9, err_msg=key)
    assert np.abs(raw[3][0] - clamped[3][0]).max() > 1e-3
    plan = _plan(np.triu(p), a, l, u)
    xc, _, yc, *_ = oracle.OracleOSQP(plan).solve(pv[None], av[None], q[None], l[None], u[None])
    a1 = oracle.adjoint_exact(plan, pv, av, l, u, xc[0], yc[0], xb, yb, clamp_slacks=True)
    a2 = oracle.adjoint_exact(plan, pv, av, l, u, xc[0], yc[0], xb, yb)
    for v1, v2 in zip(a1, a2):
        np.testing.assert_allclose(v1, v2, rtol=1e-9, atol=1e-12)
