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


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
