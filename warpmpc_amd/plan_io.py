"""Versioned on-disk format of QDLDL and OSQP plans, and a plan cache.

The reference rebuilds its ``QDLDLPlan`` (``warpmpc/jax_qdldl/types.py:11-53``)
and ``FixedOSQPPlan`` (``warpmpc/jax_osqp/types.py:86-132``) in Python every
time a solver is compiled.  Planning (AMD ordering, elimination tree, row
programs, Ruiz scaling, the adjoint pattern) is host work of seconds for the
larger patterns, so plans can be written once and loaded later:

* ``save_qdldl_plan`` / ``load_qdldl_plan`` and ``save_osqp_plan`` /
  ``load_osqp_plan`` write and read one ``.npz`` file (numpy's zip of ``.npy``
  arrays, read with ``allow_pickle=False``, so loading executes nothing from
  the file).  Every array field is stored under its dotted field path; scalar
  fields, nested plans, sparse matrices and the settings are described by a
  JSON header stored as a byte array.  The header carries the format name, its
  version and a SHA-256 over every stored array; a file of another format, a
  newer version or with a mismatching digest is rejected.
* ``cached_osqp_plan`` keys a plan by a digest of its inputs (P and A pattern
  and values, bounds, settings, derivative flag) and keeps it in a cache
  directory (``WARPMPC_AMD_PLAN_CACHE``, default ``~/.cache/warpmpc_amd/plans``).

Loaded plans compare equal, field by field, to the plans that were saved; an
externally supplied permutation survives the round trip as ``p``.
"""

from __future__ import annotations

import dataclasses
import hashlib
import io
import json
import os
import pathlib

import numpy as np
import scipy.sparse as sp

FORMAT = "warpmpc_amd-plan"
VERSION = 1


def _types() -> dict:
    # imported on use: the osqp package re-exports this module's functions
    from .osqp.types import FixedOSQPPlan, OSQPDerivativePlan, OSQPSettings
    from .qdldl.types import QDLDLPlan

    return {"QDLDLPlan": QDLDLPlan, "OSQPDerivativePlan": OSQPDerivativePlan,
            "FixedOSQPPlan": FixedOSQPPlan, "OSQPSettings": OSQPSettings}


def _encode(obj, prefix: str, arrays: dict) -> dict:
    """JSON description of a dataclass; its arrays go into ``arrays``."""
    name = type(obj).__name__
    if name not in _types():
        raise TypeError(f"cannot serialise {name}")
    desc = {"type": name, "fields": {}}
    for f in dataclasses.fields(obj):
        v = getattr(obj, f.name)
        key = f"{prefix}{f.name}"
        if v is None:
            desc["fields"][f.name] = {"kind": "none"}
        elif isinstance(v, np.ndarray):
            arrays[key] = np.ascontiguousarray(v)
            desc["fields"][f.name] = {"kind": "array"}
        elif sp.issparse(v):
            m = sp.csc_matrix(v)
            arrays[f"{key}.indptr"] = m.indptr
            arrays[f"{key}.indices"] = m.indices
            arrays[f"{key}.data"] = m.data
            desc["fields"][f.name] = {"kind": "csc", "shape": list(m.shape)}
        elif dataclasses.is_dataclass(v):
            desc["fields"][f.name] = {"kind": "dataclass", "value": _encode(v, f"{key}.", arrays)}
        elif isinstance(v, (bool, np.bool_)):
            desc["fields"][f.name] = {"kind": "bool", "value": bool(v)}
        elif isinstance(v, (int, np.integer)):
            desc["fields"][f.name] = {"kind": "int", "value": int(v)}
        elif isinstance(v, (float, np.floating)):
            desc["fields"][f.name] = {"kind": "float", "value": float(v).hex()}
        elif isinstance(v, str):
            desc["fields"][f.name] = {"kind": "str", "value": v}
        else:
            raise TypeError(f"field {key} of type {type(v).__name__} cannot be serialised")
    return desc


def _decode(desc: dict, prefix: str, arrays) -> object:
    cls = _types().get(desc.get("type"))
    if cls is None:
        raise ValueError(f"unknown plan type {desc.get('type')!r}")
    names = {f.name for f in dataclasses.fields(cls)}
    if set(desc["fields"]) != names:
        raise ValueError(f"{cls.__name__} fields in the file do not match this version: "
                         f"{sorted(set(desc['fields']) ^ names)}")
    kw = {}
    for name, fd in desc["fields"].items():
        key = f"{prefix}{name}"
        kind = fd["kind"]
        if kind == "none":
            kw[name] = None
        elif kind == "array":
            kw[name] = arrays[key]
        elif kind == "csc":
            kw[name] = sp.csc_matrix((arrays[f"{key}.data"], arrays[f"{key}.indices"],
                                      arrays[f"{key}.indptr"]), shape=tuple(fd["shape"]))
        elif kind == "dataclass":
            kw[name] = _decode(fd["value"], f"{key}.", arrays)
        elif kind == "bool":
            kw[name] = bool(fd["value"])
        elif kind == "int":
            kw[name] = int(fd["value"])
        elif kind == "float":
            kw[name] = float.fromhex(fd["value"])
        elif kind == "str":
            kw[name] = str(fd["value"])
        else:
            raise ValueError(f"unknown field kind {kind!r} for {key}")
    return cls(**kw)


def _digest(arrays: dict) -> str:
    h = hashlib.sha256()
    for key in sorted(arrays):
        a = np.ascontiguousarray(arrays[key])
        h.update(key.encode())
        h.update(str(a.dtype).encode())
        h.update(str(a.shape).encode())
        h.update(a.tobytes())
    return h.hexdigest()


def _save(obj, path) -> pathlib.Path:
    arrays: dict = {}
    desc = _encode(obj, "", arrays)
    header = {"format": FORMAT, "version": VERSION, "plan": desc, "sha256": _digest(arrays)}
    path = pathlib.Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    buf = io.BytesIO()
    np.savez(buf, __header__=np.frombuffer(json.dumps(header).encode(), dtype=np.uint8), **arrays)
    tmp = path.with_name(path.name + f".tmp{os.getpid()}")
    tmp.write_bytes(buf.getvalue())
    os.replace(tmp, path)
    return path


def _load(path, expect: type):
    with np.load(pathlib.Path(path), allow_pickle=False) as f:
        if "__header__" not in f.files:
            raise ValueError(f"{path} is not a {FORMAT} file")
        header = json.loads(f["__header__"].tobytes().decode())
        if header.get("format") != FORMAT:
            raise ValueError(f"{path} is not a {FORMAT} file")
        if int(header.get("version", -1)) > VERSION:
            raise ValueError(f"{path} has format version {header['version']}, newer than {VERSION}")
        arrays = {k: f[k] for k in f.files if k != "__header__"}
    if _digest(arrays) != header.get("sha256"):
        raise ValueError(f"{path}: array digest mismatch (file damaged or edited)")
    obj = _decode(header["plan"], "", arrays)
    if not isinstance(obj, expect):
        raise ValueError(f"{path} holds a {type(obj).__name__}, not a {expect.__name__}")
    return obj


def save_qdldl_plan(plan, path) -> pathlib.Path:
    """Writes a ``QDLDLPlan``."""
    return _save(plan, path)


def load_qdldl_plan(path):
    """Reads a ``QDLDLPlan`` written by ``save_qdldl_plan``."""
    return _load(path, _types()["QDLDLPlan"])


def save_osqp_plan(plan, path) -> pathlib.Path:
    """Writes a ``FixedOSQPPlan`` (with its QDLDL and derivative plans)."""
    return _save(plan, path)


def load_osqp_plan(path):
    """Reads a ``FixedOSQPPlan`` written by ``save_osqp_plan``."""
    return _load(path, _types()["FixedOSQPPlan"])


def plan_cache_dir() -> pathlib.Path:
    env = os.environ.get("WARPMPC_AMD_PLAN_CACHE")
    return pathlib.Path(env) if env else pathlib.Path.home() / ".cache" / "warpmpc_amd" / "plans"


def osqp_plan_key(p_matrix, a_matrix, l, u, settings, derivatives: bool = False) -> str:
    """Digest of everything ``build_osqp_plan`` reads."""
    h = hashlib.sha256()
    h.update(f"{FORMAT}/{VERSION}".encode())
    for mat in (p_matrix, a_matrix):
        m = sp.csc_matrix(mat)
        m.sort_indices()
        for a in (np.asarray(m.shape), m.indptr, m.indices, m.data.astype(np.float64)):
            h.update(np.ascontiguousarray(a).tobytes())
    for v in (l, u):
        h.update(np.ascontiguousarray(np.asarray(v, dtype=np.float64)).tobytes())
    h.update(json.dumps(dataclasses.asdict(settings), sort_keys=True).encode())
    h.update(b"d" if derivatives else b"-")
    return h.hexdigest()[:32]


def cached_osqp_plan(p_matrix, a_matrix, l, u, settings=None, *, derivatives: bool = False,
                     cache_dir=None):
    """``build_osqp_plan`` through the on-disk cache."""
    from .osqp.planning import build_osqp_plan

    settings = settings or _types()["OSQPSettings"]()
    path = pathlib.Path(cache_dir or plan_cache_dir()) / (
        osqp_plan_key(p_matrix, a_matrix, l, u, settings, derivatives) + ".npz")
    if path.exists():
        try:
            return load_osqp_plan(path)
        except ValueError:
            pass  # stale or damaged entry: rebuild and overwrite
    plan = build_osqp_plan(p_matrix, a_matrix, l, u, settings, derivatives=derivatives)
    save_osqp_plan(plan, path)
    return plan


__all__ = [
    "FORMAT",
    "VERSION",
    "cached_osqp_plan",
    "load_osqp_plan",
    "load_qdldl_plan",
    "osqp_plan_key",
    "plan_cache_dir",
    "save_osqp_plan",
    "save_qdldl_plan",
]
