"""Heterogeneous-batch workload definitions and the MPC shift maps (CPU only)."""

from __future__ import annotations

import numpy as np

from warpmpc_amd.problems import HETERO_SHAPES, hetero_counts, hetero_mpc_patterns, mass_chain_dynamics, mpc_qp, \
    mpc_shift_maps


def test_nine_patterns_have_the_stated_sizes():
    pats = hetero_mpc_patterns()
    assert [(p.nx, p.nu, p.horizon) for p in pats] == list(HETERO_SHAPES)
    for p in pats:
        assert p.n == (p.horizon + 1) * p.nx + p.horizon * p.nu
        assert p.m == (p.horizon + 1) * p.nx + p.n
    assert (pats[0].n, pats[0].m) == (64, 108) and (pats[-1].n, pats[-1].m) == (1304, 2288)
    # every input moves the chain (no zero column in B)
    for p in pats:
        assert np.all(np.abs(p.bd).sum(axis=0) > 0)


def test_counts_draw_every_pattern_uniformly():
    c = hetero_counts(200_000)
    assert c.sum() == 200_000 and c.size == 9
    assert np.all(np.abs(c - 200_000 / 9) < 5 * np.sqrt(200_000 / 9))


def test_shift_maps_advance_one_stage():
    ad, bd = mass_chain_dynamics(2, 2)
    prob = mpc_qp(ad, bd, 3, np.ones(4), np.ones(2))
    nx, nu, n_h = 4, 2, 3
    var, con = mpc_shift_maps(prob)
    # a solution whose entries name themselves: stage k of x holds 10k + i, of u 100 + 10k + j
    xs = np.array([10 * k + i for k in range(n_h + 1) for i in range(nx)], dtype=float)
    us = np.array([100 + 10 * k + j for k in range(n_h) for j in range(nu)], dtype=float)
    z = np.concatenate([xs, us])
    shifted = z[var]
    np.testing.assert_array_equal(shifted[:nx], xs[nx:2 * nx])                  # x_0 <- x_1
    np.testing.assert_array_equal(shifted[n_h * nx:(n_h + 1) * nx], xs[-nx:])   # x_N <- x_N
    np.testing.assert_array_equal(shifted[(n_h + 1) * nx:(n_h + 1) * nx + nu], us[nu:2 * nu])  # u_0 <- u_1
    np.testing.assert_array_equal(shifted[-nu:], us[-nu:])                     # u_{N-1} <- u_{N-1}
    # constraint rows: dynamics blocks follow the states, identity rows the variables
    c = np.concatenate([xs, z])
    np.testing.assert_array_equal(c[con][: (n_h + 1) * nx], shifted[: (n_h + 1) * nx])
    np.testing.assert_array_equal(c[con][(n_h + 1) * nx:], shifted)
    assert var.dtype == np.int32 and con.dtype == np.int32
    assert var.max() < prob.n and con.max() < prob.m
