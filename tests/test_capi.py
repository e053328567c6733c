"""The C-ABI library loads and exports every entry point include/warpmpc_hip.h declares.

CPU only: no compute call is made; argument validation that fails before any
HIP call is exercised (null pointers give WMPC_ERR_INVALID and an error text).
"""

from __future__ import annotations

import ctypes
import pathlib
import re

import pytest

from warpmpc_amd import _lib

ROOT = pathlib.Path(__file__).resolve().parents[1]


def _declared_symbols() -> set[str]:
    names = set()
    for header in (ROOT / "include").glob("*.h"):
        text = header.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(wmpc_\w+)\s*\(", text))
    return names


@pytest.fixture(scope="module")
def lib():
    if not _lib.LIB_PATH.exists():
        pytest.fail(f"{_lib.LIB_PATH} is not built (python -m warpmpc_amd.build)")
    return _lib.load()


def test_header_declares_entry_points():
    names = _declared_symbols()
    assert {"wmpc_qdldl_factor_f32", "wmpc_qdldl_solve_f32", "wmpc_osqp_solve_f32"} <= names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in sorted(_declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert _declared_symbols() <= set(_lib.SIGNATURES)


def test_null_arguments_are_rejected_without_a_gpu(lib):
    out = ctypes.c_void_p()
    rc = lib.wmpc_qdldl_plan_create(None, ctypes.byref(out))
    assert rc == 1
    assert b"null" in lib.wmpc_last_error()
    rc = lib.wmpc_qdldl_factor_f32(None, None, None, None, None, 1, 1, None)
    assert rc == 1
    assert lib.wmpc_qdldl_plan_destroy(None) == 0
    assert lib.wmpc_version().startswith(b"warpmpc_hip")
