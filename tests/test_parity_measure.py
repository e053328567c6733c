"""The GPU parity tests' error measure on known answers (CPU)."""

from __future__ import annotations

import numpy as np

from parity_measure import adjoint_errors, osqp_errors, rel_err


def test_norm_wise_relative_per_row():
    ref = np.array([[1e-3, -2e-3], [10.0, 0.0]])
    v = ref + np.array([[1e-7, 0.0], [0.0, 1e-3]])
    # row 0: 1e-7 / 2e-3 = 5e-5; row 1: 1e-3 / 10 = 1e-4
    assert np.isclose(rel_err(v, ref), 1e-4)
    assert np.isclose(rel_err(v[:1], ref[:1]), 5e-5)


def test_zero_rows_use_a_floor_from_the_batch_scale():
    ref = np.array([[0.0, 0.0], [2.0, 1.0], [4.0, 0.0]])
    v = ref.copy()
    v[0, 1] = 1e-6
    # median row scale 2 -> floor 2e-3 -> 1e-6 / 2e-3
    assert np.isclose(rel_err(v, ref), 5e-4)


def test_scalar_outputs_and_empty_width():
    assert np.isclose(rel_err(np.array([1.0001, 2.0]), np.array([1.0, 2.0])), 1e-4)
    assert rel_err(np.zeros((3, 0)), np.zeros((3, 0))) == 0.0


def test_osqp_vectors_floored_at_the_instance_scale():
    x = np.array([[2.0, 1.0], [1.0, 0.0]])
    z = np.array([[1.0], [0.5]])
    y = np.array([[0.0], [1e-6]])
    ref = (x, z, y)
    out = (x, z, y + np.array([[1e-6], [1e-6]]))
    # instance 0: 1e-6 / (1e-2 * 2); instance 1: 1e-6 / (1e-2 * 1)
    assert np.isclose(osqp_errors(out, ref)["y"], 1e-4)


def test_residuals_relative_to_iterate_scale():
    x = np.array([[3.0, 1.0]])
    z = np.array([[0.5]])
    y = np.array([[-7.0]])
    ref = (x, z, y, np.array([1e-4]), np.array([2e-4]), np.array([5.0]))
    out = (x, z, y, np.array([1e-4 + 7e-7]), np.array([2e-4]), np.array([5.0]))
    errs = osqp_errors(out, ref, q=np.array([[1.0, 0.0]]))
    assert np.isclose(errs["prim_res"], 1e-7)
    assert errs["x"] == errs["dual_res"] == errs["obj_val"] == 0.0


def test_objective_relative_to_its_terms():
    # q'x = -3, obj = 0.01: 0.5 x'Px = 3.01, term scale 6.01
    x = np.array([[3.0, 0.0]])
    ref = (x, x[:, :1], x[:, :1], np.zeros(1), np.zeros(1), np.array([0.01]))
    out = ref[:5] + (np.array([0.01 + 6.01e-4]),)
    assert np.isclose(osqp_errors(out, ref, q=np.array([[-1.0, 5.0]]))["obj_val"], 1e-4)


def test_adjoint_outputs_floored_at_the_instance_gradient_scale():
    dq = np.array([[1e-12, 0.0], [0.5, 0.2]])
    du = np.array([[1.0], [0.1]])
    got_dq = dq + np.array([[1e-7, 0.0], [0.0, 0.0]])
    # instance 0: dq ~ 0, du = 1 -> floor 1e-2 -> 1e-7 / 1e-2
    errs = adjoint_errors((got_dq, du), (dq, du), ("dq", "du"))
    assert np.isclose(errs["dq"], 1e-5) and errs["du"] == 0.0
