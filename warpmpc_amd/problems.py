"""Synthetic QP families used by the tests and the benchmark.

The QP layout is the OSQP-example MPC layout of the reference's fixture
``benchmarks/problems/linear_mpc.py:109-132``: variables
``z = [x_0 .. x_N, u_0 .. u_{N-1}]``, constraints ``A = [A_eq; I]`` with the
dynamics rows ``-x_{k+1} + Ad x_k + Bd u_k = 0`` and the first block fixing
``x_0`` through ``l = u = -x0`` (``linear_mpc.py:174-175``).

Plants:

* ``mass_chain`` — ``n_masses`` unit masses in a line, joined by springs of
  stiffness ``k`` to each other and to a wall at either end, no damping.  The
  ``nu = 4`` inputs are a force on the first mass, tensions between masses
  (1, 2) and (n-3, n-2), and a force on the last mass (the oscillating-masses
  benchmark of Wang & Boyd, "Fast model predictive control using online
  optimization", 2010, with one more actuator).  Discretised by zero-order
  hold through the augmented matrix exponential.
* ``cartpole`` — cart-pole linearised at the upright equilibrium, explicit
  Euler.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sp

OSQP_INFTY = 1e30


@dataclass(frozen=True)
class LinearMPC:
    p_matrix: sp.csc_matrix
    q: np.ndarray
    a_matrix: sp.csc_matrix
    l: np.ndarray
    u: np.ndarray
    ad: np.ndarray
    bd: np.ndarray
    nx: int
    nu: int
    horizon: int

    @property
    def n(self) -> int:
        return int(self.p_matrix.shape[0])

    @property
    def m(self) -> int:
        return int(self.a_matrix.shape[0])


def zoh(a_c: np.ndarray, b_c: np.ndarray, dt: float) -> tuple[np.ndarray, np.ndarray]:
    """Zero-order-hold discretisation via ``expm([[A, B], [0, 0]] dt)``."""
    nx, nu = b_c.shape
    aug = np.zeros((nx + nu, nx + nu))
    aug[:nx, :nx] = a_c
    aug[:nx, nx:] = b_c
    e = sla.expm(aug * dt)
    return e[:nx, :nx], e[:nx, nx:]


def mass_chain_dynamics(n_masses: int = 6, nu: int = 4, springs=None, dt: float = 0.5):
    """Continuous chain of unit masses with wall springs at both ends, ZOH-discretised."""
    if springs is None:
        springs = np.ones(n_masses + 1)
    springs = np.asarray(springs, dtype=float)
    if springs.shape != (n_masses + 1,):
        raise ValueError("need n_masses + 1 spring constants")
    stiff = np.zeros((n_masses, n_masses))
    for i in range(n_masses):
        stiff[i, i] = springs[i] + springs[i + 1]
        if i + 1 < n_masses:
            stiff[i, i + 1] = -springs[i + 1]
            stiff[i + 1, i] = -springs[i + 1]
    force = np.zeros((n_masses, nu))
    if nu >= 1:
        force[0, 0] = 1.0
    if nu >= 2:
        if n_masses >= 3:  # pushes mass 1 against mass 2
            force[1, 1] += 1.0
            force[2, 1] -= 1.0
        else:
            force[n_masses - 1, 1] += 1.0
    if nu >= 3:
        force[max(n_masses - 3, 0), 2] += 1.0
        force[max(n_masses - 2, 0), 2] -= 1.0
    if nu >= 4:
        force[n_masses - 1, 3] = 1.0
    for extra in range(4, nu):
        force[extra % n_masses, extra] = 1.0
    nx = 2 * n_masses
    a_c = np.zeros((nx, nx))
    a_c[:n_masses, n_masses:] = np.eye(n_masses)
    a_c[n_masses:, :n_masses] = -stiff
    b_c = np.zeros((nx, nu))
    b_c[n_masses:, :] = force
    return zoh(a_c, b_c, dt)


def mpc_qp(
    ad: np.ndarray,
    bd: np.ndarray,
    horizon: int,
    q_diag: np.ndarray,
    r_diag: np.ndarray,
    qn_diag: np.ndarray | None = None,
    x_bound: float | np.ndarray = 4.0,
    u_bound: float | np.ndarray = 0.5,
    x0: np.ndarray | None = None,
) -> LinearMPC:
    """OSQP-example MPC QP (``linear_mpc.py:109-132`` layout) with box bounds."""
    nx, nu = bd.shape
    n_h = int(horizon)
    qn_diag = q_diag if qn_diag is None else qn_diag
    p_matrix = sp.block_diag(
        [
            sp.kron(sp.eye(n_h), sp.diags(q_diag)),
            sp.diags(qn_diag),
            sp.kron(sp.eye(n_h), sp.diags(r_diag)),
        ],
        format="csc",
    )
    # kron of small diagonal blocks can store their off-diagonal zeros
    p_matrix.eliminate_zeros()
    p_matrix.sort_indices()
    ax = sp.kron(sp.eye(n_h + 1), -sp.eye(nx)) + sp.kron(sp.eye(n_h + 1, k=-1), sp.csc_matrix(ad))
    bu = sp.kron(sp.vstack([sp.csc_matrix((1, n_h)), sp.eye(n_h)]), sp.csc_matrix(bd))
    aeq = sp.hstack([ax, bu])
    n = (n_h + 1) * nx + n_h * nu
    a_matrix = sp.vstack([aeq, sp.eye(n)], format="csc")
    a_matrix.sum_duplicates()
    a_matrix.sort_indices()
    x0 = np.zeros(nx) if x0 is None else np.asarray(x0, dtype=float)
    xb = np.broadcast_to(np.asarray(x_bound, dtype=float), (nx,))
    ub = np.broadcast_to(np.asarray(u_bound, dtype=float), (nu,))
    leq = np.hstack([-x0, np.zeros(n_h * nx)])
    lineq = np.hstack([np.tile(-xb, n_h + 1), np.tile(-ub, n_h)])
    uineq = np.hstack([np.tile(xb, n_h + 1), np.tile(ub, n_h)])
    return LinearMPC(
        p_matrix=p_matrix.tocsc(),
        q=np.zeros(n),
        a_matrix=a_matrix,
        l=np.hstack([leq, lineq]),
        u=np.hstack([leq.copy(), uineq]),
        ad=np.asarray(ad),
        bd=np.asarray(bd),
        nx=nx,
        nu=nu,
        horizon=n_h,
    )


def mass_chain_mpc(horizon: int = 10, n_masses: int = 6, nu: int = 4, dt: float = 0.5,
                   springs=None, q_diag=None, r_diag=None, x_bound=4.0, u_bound=0.5) -> LinearMPC:
    """Mass-chain linear MPC: Q = I, R = I, terminal = Q, |x| <= 4, |u| <= 0.5 by default."""
    ad, bd = mass_chain_dynamics(n_masses, nu, springs=springs, dt=dt)
    nx = 2 * n_masses
    q_diag = np.ones(nx) if q_diag is None else np.asarray(q_diag, dtype=float)
    r_diag = np.ones(nu) if r_diag is None else np.asarray(r_diag, dtype=float)
    return mpc_qp(ad, bd, horizon, q_diag, r_diag, x_bound=x_bound, u_bound=u_bound)


def cartpole_dynamics(m_cart: float = 1.0, m_pole: float = 0.1, length: float = 0.5,
                      g: float = 9.81, dt: float = 0.02) -> tuple[np.ndarray, np.ndarray]:
    """Cart-pole linearised at the upright equilibrium; state [p, theta, v, omega]; Euler.

    Point-mass pole of length ``length`` on a frictionless cart, force input.
    """
    total = m_cart + m_pole
    a_c = np.zeros((4, 4))
    a_c[0, 2] = 1.0
    a_c[1, 3] = 1.0
    a_c[2, 1] = -m_pole * g / m_cart
    a_c[3, 1] = total * g / (m_cart * length)
    b_c = np.zeros((4, 1))
    b_c[2, 0] = 1.0 / m_cart
    b_c[3, 0] = -1.0 / (m_cart * length)
    return np.eye(4) + dt * a_c, dt * b_c


def batched_mpc_data(problem: LinearMPC, batch: int, *, seed: int = 0, x0_low=-1.0, x0_high=1.0,
                     dtype=np.float64):
    """Batch of (P, A, q, l, u) with shared P/A (batch dim 1) and x0 ~ U[low, high]."""
    rng = np.random.default_rng(seed)
    p_values = np.asarray(problem.p_matrix.data, dtype=dtype)[None, :]
    a_values = np.asarray(problem.a_matrix.data, dtype=dtype)[None, :]
    q = np.broadcast_to(np.asarray(problem.q, dtype=dtype), (batch, problem.n)).copy()
    l = np.broadcast_to(np.asarray(problem.l, dtype=dtype), (batch, problem.m)).copy()
    u = np.broadcast_to(np.asarray(problem.u, dtype=dtype), (batch, problem.m)).copy()
    x0 = rng.uniform(x0_low, x0_high, size=(batch, problem.nx)).astype(dtype)
    l[:, : problem.nx] = -x0
    u[:, : problem.nx] = -x0
    return p_values, a_values, q, l, u


HETERO_SHAPES = tuple((nx, nu, n_h) for nx, nu in ((4, 2), (12, 4), (24, 8)) for n_h in (10, 20, 40))


def hetero_mpc_patterns(seed: int = 2) -> list[LinearMPC]:
    """The nine patterns of the heterogeneous-batch workload.

    Mass chains with (nx, nu) in {(4, 2), (12, 4), (24, 8)} and horizon N in
    {10, 20, 40}; per pattern, springs k ~ U[0.5, 2], Q = diag(U[0.1, 10]) and
    R = diag(U[0.1, 1]) are drawn once; ZOH dt = 0.5, |x| <= 4, |u| <= 0.5.
    """
    rng = np.random.default_rng(seed)
    out = []
    for nx, nu, n_h in HETERO_SHAPES:
        nm = nx // 2
        ad, bd = mass_chain_dynamics(nm, nu, springs=rng.uniform(0.5, 2.0, nm + 1), dt=0.5)
        out.append(mpc_qp(ad, bd, n_h, rng.uniform(0.1, 10.0, nx), rng.uniform(0.1, 1.0, nu),
                          x_bound=4.0, u_bound=0.5))
    return out


def hetero_counts(total: int, patterns: int = 9, seed: int = 2) -> np.ndarray:
    """Instances per pattern when `total` instances draw their pattern uniformly."""
    rng = np.random.default_rng(seed + 1000)
    return np.bincount(rng.integers(0, patterns, size=total), minlength=patterns)


def mpc_shift_maps(problem: LinearMPC) -> tuple[np.ndarray, np.ndarray]:
    """Source rows of the one-stage shift of an MPC solution (the previous
    solution advanced one stage, the last stage repeated), for the mpc_qp
    layout: variables [x_0..x_N, u_0..u_{N-1}], constraints [dynamics blocks
    0..N (block k holds x_k's equation); identity rows of the variables].
    Returns (variable map [n], constraint map [m]): shifted[i] = previous[map[i]].
    """
    nx, nu, n_h = problem.nx, problem.nu, problem.horizon
    xs = np.minimum(np.arange(n_h + 1) + 1, n_h)
    us = np.minimum(np.arange(n_h) + 1, n_h - 1)
    xmap = (xs[:, None] * nx + np.arange(nx)).ravel()
    umap = (n_h + 1) * nx + (us[:, None] * nu + np.arange(nu)).ravel()
    var = np.concatenate([xmap, umap]).astype(np.int32)
    con = np.concatenate([xmap, (n_h + 1) * nx + var]).astype(np.int32)
    return var, con


__all__ = [
    "HETERO_SHAPES",
    "LinearMPC",
    "OSQP_INFTY",
    "batched_mpc_data",
    "cartpole_dynamics",
    "hetero_counts",
    "hetero_mpc_patterns",
    "mpc_shift_maps",
    "kkt_microbench_plan",
    "kkt_microbench_values",
    "mass_chain_dynamics",
    "mass_chain_mpc",
    "mpc_qp",
    "zoh",
]


def kkt_microbench_plan(horizon: int = 50, n_masses: int = 6, nu: int = 4, sigma: float = 1e-6):
    """OSQP plan whose KKT ``[P + sigma I, Aᵀ; A, -diag(1/rho)]`` is the mass-chain MPC KKT."""
    from .osqp.planning import build_osqp_plan
    from .osqp.types import OSQPSettings

    prob = mass_chain_mpc(horizon=horizon, n_masses=n_masses, nu=nu)
    settings = OSQPSettings(rho=0.1, sigma=sigma, rho_is_vec=False, max_iter=25)
    return prob, build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, settings)


def kkt_microbench_values(osqp_plan, batch: int, *, seed: int = 1, rho_low: float = 1e-3,
                          rho_high: float = 1e2, device=None):
    """Per-instance KKT values ``(nnz_kkt, batch)`` and rhs ``(n_kkt, batch)``, float32.

    P diagonal = 1 + U[0, 1] (+ sigma), A = the chain values, constraint diagonal
    = -1 / rho with rho log-uniform in [rho_low, rho_high] per entry and
    instance, rhs ~ N(0, 1).  With ``device`` (a torch device) the arrays are
    generated there with a seeded torch generator; otherwise with numpy.
    """
    nnz = osqp_plan.nnz_kkt
    nk = osqp_plan.qdldl_plan.n
    base = np.asarray(osqp_plan.kkt_upper.data, dtype=np.float64)
    p_pos = np.asarray(osqp_plan.p_to_kkt)
    p_diag = np.asarray(osqp_plan.p_diag_sigma) > 0
    rho_pos = np.asarray(osqp_plan.rho_to_kkt)
    sigma = float(osqp_plan.settings.sigma)
    lo, hi = np.log(rho_low), np.log(rho_high)
    if device is None:
        rng = np.random.default_rng(seed)
        vals = np.repeat(base.astype(np.float32)[:, None], batch, axis=1)
        dpos = p_pos[p_diag]
        vals[dpos] = (1.0 + rng.uniform(0.0, 1.0, size=(dpos.size, batch)) + sigma).astype(np.float32)
        rho = np.exp(rng.uniform(lo, hi, size=(rho_pos.size, batch)))
        vals[rho_pos] = (-1.0 / rho).astype(np.float32)
        rhs = rng.standard_normal((nk, batch)).astype(np.float32)
        return vals, rhs
    import torch

    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    vals = torch.as_tensor(base.astype(np.float32), device=device)[:, None].repeat(1, batch)
    dpos = torch.as_tensor(p_pos[p_diag], device=device, dtype=torch.long)
    vals[dpos] = 1.0 + torch.rand((dpos.numel(), batch), generator=g, device=device) + sigma
    rp = torch.as_tensor(rho_pos, device=device, dtype=torch.long)
    u01 = torch.rand((rp.numel(), batch), generator=g, device=device, dtype=torch.float64)
    vals[rp] = (-1.0 / torch.exp(lo + (hi - lo) * u01)).to(torch.float32)
    rhs = torch.randn((nk, batch), generator=g, device=device)
    del nnz
    return vals, rhs
