"""Nonlinear MPC problems for the SQP: a two-stage test problem and the quadrotor.

``tiny_problem`` is the two-stage problem of the reference's SQP tests
(``tests/test_jax_sqp.py:62-93``) written as torch stage functions.

``quadrotor_problem`` is the rigid-body quadrotor of the nonlinear SQP
workload: mass 0.027 kg, inertia diag(1.4e-5, 1.4e-5, 2.17e-5) kg m², arm
0.046 m, four motor thrusts in [0, 0.15] N, RK4 with dt = 0.02 s, horizon 25.
State x = [position (3), roll/pitch/yaw (3), velocity (3), body rates (3)];
the rotors sit in a '+' arrangement (roll torque arm (T2 - T4), pitch torque
arm (T3 - T1)) and the yaw torque is 0.006 m x (T1 - T2 + T3 - T4).  Stage k
holds z_k = [x_k, u_k]; the stage cost tracks a target position (stage
parameter) and the hover thrust; the stage constraints are the RK4 defect
x_{k+1} - F(x_k, u_k) = 0 and 0 <= u_k <= 0.15, and the first stage also
pins x_0 to its parameter.
"""

from __future__ import annotations

import numpy as np

from .sparse_mpc import SparseMPCProblem
from .stage import TorchStageFunction, device_const


def _torch():
    import torch

    return torch


_TINY_LOWER = np.array([-1.0, 0.0])
_TINY_UPPER = np.array([1.0, 0.0])


def tiny_problem() -> SparseMPCProblem:
    th = _torch()

    def first(z, zn, p):
        cost = 0.5 * ((z[..., 0] - p[..., 0]) ** 2 + 2.0 * z[..., 1] ** 2) + 0.25 * zn[..., 0] ** 2
        g = th.stack([z[..., 0] + zn[..., 0] ** 2, z[..., 1] + zn[..., 1]], dim=-1)
        lower = th.broadcast_to(device_const(_TINY_LOWER, z.dtype, z.device), g.shape)
        upper = th.broadcast_to(device_const(_TINY_UPPER, z.dtype, z.device), g.shape)
        return cost, g, lower, upper

    def terminal(z, p):
        cost = 0.5 * ((z[..., 0] - p[..., 0]) ** 2 + z[..., 1] ** 2)
        g = z[..., :1]
        return cost, g, th.full_like(g, -2.0), th.full_like(g, 2.0)

    f = TorchStageFunction.from_function(first, z_dim=2, next_z_dim=2, param_dim=1, name="tiny_first")
    t = TorchStageFunction.from_function(terminal, z_dim=2, param_dim=1, name="tiny_terminal")
    return SparseMPCProblem.from_stage_functions(horizon=1, first=f, intermediate=f, terminal=t)


QUAD_MASS = 0.027
QUAD_INERTIA = (1.4e-5, 1.4e-5, 2.17e-5)
QUAD_ARM = 0.046
QUAD_YAW_RATIO = 0.006
QUAD_T_MAX = 0.15
QUAD_DT = 0.02
QUAD_G = 9.81
QUAD_NX, QUAD_NU = 12, 4
QUAD_HOVER = QUAD_MASS * QUAD_G / 4.0


def quadrotor_dynamics(x, u):
    """Continuous-time rigid-body quadrotor, any leading batch shape."""
    th = _torch()
    phi, theta, psi = x[..., 3], x[..., 4], x[..., 5]
    v = x[..., 6:9]
    wx, wy, wz = x[..., 9], x[..., 10], x[..., 11]
    t1, t2, t3, t4 = u[..., 0], u[..., 1], u[..., 2], u[..., 3]
    sp, cp = th.sin(phi), th.cos(phi)
    st, ct = th.sin(theta), th.cos(theta)
    ss, cs = th.sin(psi), th.cos(psi)
    tt = st / ct
    # Euler-angle rates from body rates
    dphi = wx + sp * tt * wy + cp * tt * wz
    dtheta = cp * wy - sp * wz
    dpsi = (sp * wy + cp * wz) / ct
    # thrust along the body z axis, rotated to the world frame (R = Rz Ry Rx)
    f = (t1 + t2 + t3 + t4) / QUAD_MASS
    ax = f * (cs * st * cp + ss * sp)
    ay = f * (ss * st * cp - cs * sp)
    az = f * (ct * cp) - QUAD_G
    jx, jy, jz = QUAD_INERTIA
    tx = QUAD_ARM * (t2 - t4)
    ty = QUAD_ARM * (t3 - t1)
    tz = QUAD_YAW_RATIO * (t1 - t2 + t3 - t4)
    dwx = (tx - (jz - jy) * wy * wz) / jx
    dwy = (ty - (jx - jz) * wz * wx) / jy
    dwz = (tz - (jy - jx) * wx * wy) / jz
    return th.cat([v, th.stack([dphi, dtheta, dpsi, ax, ay, az, dwx, dwy, dwz], dim=-1)], dim=-1)


def quadrotor_rk4(x, u, dt: float = QUAD_DT):
    k1 = quadrotor_dynamics(x, u)
    k2 = quadrotor_dynamics(x + 0.5 * dt * k1, u)
    k3 = quadrotor_dynamics(x + 0.5 * dt * k2, u)
    k4 = quadrotor_dynamics(x + dt * k3, u)
    return x + (dt / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


# constants as host arrays, moved to the device once (device_const) so the
# stage functions make no host-to-device copy per call and can be captured
_Q = np.array([10.0, 10.0, 10.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0.1, 0.1, 0.1])
_R = 100.0
_QN = 10.0


def quadrotor_problem(horizon: int = 25) -> SparseMPCProblem:
    """Quadrotor MPC; params: first stage [x0 (12), target (3)], others [target (3)]."""
    th = _torch()
    nx, nu = QUAD_NX, QUAD_NU

    def tracking(x, u, target):
        q = device_const(_Q, x.dtype, x.device)
        e = th.cat([x[..., :3] - target, x[..., 3:]], dim=-1)
        du = u - QUAD_HOVER
        return 0.5 * (q * e * e).sum(dim=-1) + 0.5 * _R * (du * du).sum(dim=-1)

    def transition(z, zn, p, x0=None):
        x, u = z[..., :nx], z[..., nx:]
        defect = zn[..., :nx] - quadrotor_rk4(x, u)
        parts = [defect, u]
        lo = [th.zeros_like(defect), th.zeros_like(u)]
        up = [th.zeros_like(defect), th.full_like(u, QUAD_T_MAX)]
        if x0 is not None:
            parts.insert(0, x - x0)
            lo.insert(0, th.zeros_like(x))
            up.insert(0, th.zeros_like(x))
        return th.cat(parts, dim=-1), th.cat(lo, dim=-1), th.cat(up, dim=-1)

    def first(z, zn, p):
        g, lo, up = transition(z, zn, p, x0=p[..., :nx])
        return tracking(z[..., :nx], z[..., nx:], p[..., nx:]), g, lo, up

    def middle(z, zn, p):
        g, lo, up = transition(z, zn, p)
        return tracking(z[..., :nx], z[..., nx:], p), g, lo, up

    def terminal(z, p):
        x, u = z[..., :nx], z[..., nx:]
        cost = _QN * tracking(x, u, p)
        return cost, u, th.zeros_like(u), th.full_like(u, QUAD_T_MAX)

    nz = nx + nu
    f = TorchStageFunction.from_function(first, z_dim=nz, next_z_dim=nz, param_dim=nx + 3, name="quad_first")
    m = TorchStageFunction.from_function(middle, z_dim=nz, next_z_dim=nz, param_dim=3, name="quad_stage")
    t = TorchStageFunction.from_function(terminal, z_dim=nz, param_dim=3, name="quad_terminal")
    return SparseMPCProblem.from_stage_functions(horizon=horizon, first=f, intermediate=m, terminal=t)


def quadrotor_initial_guess_and_params(batch: int, horizon: int = 25, seed: int = 4):
    """x0: position ±1 m, roll/pitch/yaw ±30°, velocity ±1 m/s, rates ±1 rad/s;
    target: the origin; initial guess: hover at x0 over the horizon."""
    rng = np.random.default_rng(seed)
    deg30 = np.deg2rad(30.0)
    x0 = np.concatenate([rng.uniform(-1, 1, (batch, 3)), rng.uniform(-deg30, deg30, (batch, 3)),
                         rng.uniform(-1, 1, (batch, 3)), rng.uniform(-1, 1, (batch, 3))], axis=1)
    target = np.zeros((batch, 3))
    stage = np.concatenate([x0, np.full((batch, QUAD_NU), QUAD_HOVER)], axis=1)
    z = np.tile(stage, (1, horizon + 1))
    params = np.concatenate([x0, target] + [target] * horizon, axis=1)
    return z, params


__all__ = [
    "QUAD_DT",
    "QUAD_HOVER",
    "QUAD_NU",
    "QUAD_NX",
    "QUAD_T_MAX",
    "quadrotor_dynamics",
    "quadrotor_initial_guess_and_params",
    "quadrotor_problem",
    "quadrotor_rk4",
    "tiny_problem",
]
