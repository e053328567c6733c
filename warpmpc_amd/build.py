"""Builds the HIP library ``warpmpc_amd/libwarpmpc_hip.so`` for gfx950 with hipcc.

Run ``python -m warpmpc_amd.build``.  The library is built in-tree so that it
travels with the source tree; nothing is JIT-compiled at import time.
"""

from __future__ import annotations

import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libwarpmpc_hip.so"
ARCH = os.environ.get("WMPC_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    f"--offload-arch={ARCH}",
    # keep the compiler from emitting packed-FP32 VALU forms
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not pathlib.Path(path).exists():
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def sources() -> list[pathlib.Path]:
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = False) -> pathlib.Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", str(tmp), *map(str, sources())]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def specialized_plans():
    """OSQP plans whose shared-factor solves ship with a generated
    straight-line kernel: the mass-chain linear MPC (nx=12, nu=4) at horizons
    10 and 20.  The kernel depends only on the KKT pattern, so any settings and
    data of these problems use it."""
    from .osqp import OSQPSettings, build_osqp_plan
    from .problems import mass_chain_mpc

    plans = []
    for horizon in (10, 20):
        prob = mass_chain_mpc(horizon=horizon)
        plans.append(build_osqp_plan(prob.p_matrix, prob.a_matrix, prob.l, prob.u,
                                     OSQPSettings(rho=0.1, sigma=1e-6, alpha=1.6, rho_is_vec=True)))
    return plans


def build_kernels(verbose: bool = False, force: bool = True) -> list[pathlib.Path]:
    """Generates and compiles the straight-line kernels of specialized_plans()."""
    from concurrent.futures import ThreadPoolExecutor

    from . import codegen

    plans = specialized_plans()
    with ThreadPoolExecutor(max_workers=min(4, len(plans))) as pool:
        return list(pool.map(lambda pl: codegen.build_admm_kernel(pl, force=force, verbose=verbose), plans))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    if "--no-kernels" not in sys.argv:
        print(*build_kernels(verbose="-v" in sys.argv), sep="\n")
