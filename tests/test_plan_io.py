"""On-disk plan format and the layout adapters (CPU)."""

from __future__ import annotations

import dataclasses
import json
import zipfile

import numpy as np
import pytest
import scipy.sparse as sp

from warpmpc_amd import plan_io
from warpmpc_amd.osqp import (OSQPSettings, batch_from_compiled_qdldl, build_osqp_plan, rhs_for_compiled_qdldl,
                              values_for_compiled_qdldl)
from warpmpc_amd.problems import OSQP_INFTY, cartpole_dynamics, mass_chain_mpc, mpc_qp
from warpmpc_amd.qdldl import build_qdldl_plan, load_qdldl_plan, save_qdldl_plan
from warpmpc_amd.qdldl.types import CompiledQDLDL


def _same(a, b, path="plan"):
    """Field-by-field equality of two plans (arrays bitwise, sparse by parts)."""
    assert type(a) is type(b), path
    if dataclasses.is_dataclass(a):
        for f in dataclasses.fields(a):
            _same(getattr(a, f.name), getattr(b, f.name), f"{path}.{f.name}")
    elif sp.issparse(a):
        assert a.shape == b.shape, path
        for part in ("indptr", "indices", "data"):
            x, y = getattr(a, part), getattr(b, part)
            assert x.dtype == y.dtype and np.array_equal(x, y), f"{path}.{part}"
    elif isinstance(a, np.ndarray):
        assert a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b), path
    elif isinstance(a, float):
        assert a == b or (np.isnan(a) and np.isnan(b)), path
    else:
        assert a == b, path


def _chain_plan(derivatives=False):
    prob = mass_chain_mpc(horizon=4)
    s = OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, max_iter=25, rho_is_vec=True, scaling=3)
    return prob, s, build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s, derivatives=derivatives)


def test_qdldl_plan_round_trip_keeps_an_external_permutation(tmp_path):
    rng = np.random.default_rng(0)
    n = 30
    a = sp.random(n, n, density=0.15, random_state=1)
    a = sp.triu(a + a.T + sp.eye(n) * 5, format="csc")
    p = rng.permutation(n).astype(np.int32)
    plan = build_qdldl_plan(a, p=p)
    path = save_qdldl_plan(plan, tmp_path / "k.npz")
    back = load_qdldl_plan(path)
    _same(plan, back)
    assert np.array_equal(back.p, p)


@pytest.mark.parametrize("derivatives", [False, True])
def test_osqp_plan_round_trip(tmp_path, derivatives):
    _, _, plan = _chain_plan(derivatives)
    back = plan_io.load_osqp_plan(plan_io.save_osqp_plan(plan, tmp_path / "o.npz"))
    _same(plan, back)
    assert (back.derivative_plan is None) == (not derivatives)


def test_cartpole_plan_with_infinite_bounds_round_trips(tmp_path):
    ad, bd = cartpole_dynamics()
    prob = mpc_qp(ad, bd, 5, np.ones(4), np.ones(1), x_bound=OSQP_INFTY, u_bound=10.0)
    plan = build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, OSQPSettings(rho_is_vec=True),
                           derivatives=True)
    _same(plan, plan_io.load_osqp_plan(plan_io.save_osqp_plan(plan, tmp_path / "c.npz")))


def test_damaged_foreign_and_newer_files_are_rejected(tmp_path):
    _, _, plan = _chain_plan()
    path = plan_io.save_osqp_plan(plan, tmp_path / "o.npz")
    with pytest.raises(ValueError, match="not a QDLDLPlan"):
        plan_io.load_qdldl_plan(path)
    # rewrite one array with different contents: the digest no longer matches
    with np.load(path, allow_pickle=False) as f:
        arrays = {k: f[k] for k in f.files}
    arrays["rho_vec"] = arrays["rho_vec"] * 2
    bad = tmp_path / "bad.npz"
    np.savez(bad, **arrays)
    with pytest.raises(ValueError, match="digest"):
        plan_io.load_osqp_plan(bad)
    header = json.loads(arrays["__header__"].tobytes().decode())
    header["version"] = plan_io.VERSION + 1
    arrays["rho_vec"] = arrays["rho_vec"] / 2
    arrays["__header__"] = np.frombuffer(json.dumps(header).encode(), dtype=np.uint8)
    newer = tmp_path / "newer.npz"
    np.savez(newer, **arrays)
    with pytest.raises(ValueError, match="newer"):
        plan_io.load_osqp_plan(newer)
    other = tmp_path / "other.npz"
    np.savez(other, x=np.zeros(3))
    with pytest.raises(ValueError, match="not a"):
        plan_io.load_osqp_plan(other)
    assert zipfile.is_zipfile(path)


def test_plan_cache_hits_and_keys(tmp_path):
    prob, s, plan = _chain_plan()
    first = plan_io.cached_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s, cache_dir=tmp_path)
    _same(plan, first)
    files = list(tmp_path.glob("*.npz"))
    assert len(files) == 1
    again = plan_io.cached_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s, cache_dir=tmp_path)
    _same(plan, again)
    assert list(tmp_path.glob("*.npz")) == files
    # any input that planning reads changes the key
    keys = {plan_io.osqp_plan_key(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s)}
    keys.add(plan_io.osqp_plan_key(prob.p_matrix * 2, prob.a_matrix, prob.l, prob.u, s))
    keys.add(plan_io.osqp_plan_key(prob.p_matrix, prob.a_matrix, prob.l, prob.u + 1, s))
    keys.add(plan_io.osqp_plan_key(prob.p_matrix, prob.a_matrix, prob.l, prob.u, dataclasses.replace(s, rho=0.2)))
    keys.add(plan_io.osqp_plan_key(prob.p_matrix, prob.a_matrix, prob.l, prob.u, s, derivatives=True))
    assert len(keys) == 5


def test_layout_adapters_follow_the_declared_layouts():
    def compiled(values_layout, rhs_layout):
        return CompiledQDLDL(plan=None, dtype=np.float32, factor=None, solve=None, factor_and_solve=None,
                             values_layout=values_layout, rhs_layout=rhs_layout)

    v = np.arange(12.0).reshape(3, 4)
    sym = compiled("original_symbolic", "symbolic")
    assert np.array_equal(values_for_compiled_qdldl(sym, v), v.T)
    assert np.array_equal(rhs_for_compiled_qdldl(sym, v), v.T)
    assert np.array_equal(batch_from_compiled_qdldl(sym, v.T), v)
    bat = compiled("original_batch", "batch")
    assert values_for_compiled_qdldl(bat, v) is v
    assert rhs_for_compiled_qdldl(bat, v) is v
    assert batch_from_compiled_qdldl(bat, v) is v
    with pytest.raises(ValueError, match="values layout"):
        values_for_compiled_qdldl(compiled("other", "symbolic"), v)
    with pytest.raises(ValueError, match="rhs layout"):
        rhs_for_compiled_qdldl(compiled("original_batch", "other"), v)
