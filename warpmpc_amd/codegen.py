"""Straight-line ADMM kernels generated for one KKT pattern (gfx950).

The generic ADMM kernel walks record streams of the pattern at run time: every
L entry costs a record decode (scalar ALU work, readlane, LDS window address)
on top of its multiply-add.  For a fixed pattern the whole iteration is known
at build time, so this module emits it as straight-line HIP code:

* every entry of the KKT vector is a named value (``r<c>``) that the compiler
  keeps in the lane's registers: the right-hand side before the forward
  substitution, the D⁻¹-scaled forward result between the sweeps, the next
  right-hand side after the backward substitution (each KKT position holds
  exactly one of these at any time, so the register demand is the KKT
  dimension plus the substitution band);
* the forward substitution is column (scatter) form, ``r_i -= L[q] r_c`` in
  the order of ``QDLDL_Lsolve``; the backward substitution is column (gather)
  form, ``w_c = s_c - sum L[q] w_i`` (``QDLDL_Ltsolve``);
* as soon as the backward substitution finalises KKT position c, the ADMM
  update of that entry runs (relaxation, box projection, dual update of
  solver.py:806-820, the same expressions as ``admm_pass``) and writes the
  next right-hand side of c (solver.py:808-810), so the iterate x, z, y and
  the data q, l, u are read and x, z, y written once per iteration and there
  is no stash or right-hand-side buffer in memory;
* L and D⁻¹ of the shared factor, and the per-entry constants (q scale, ρ⁻¹,
  ρ, E), are read with constant offsets from wave-uniform pointers (scalar
  loads).

The generated code depends only on the pattern: the KKT dimension split
(n variables, m constraints), the permutation and the structure of L.  Its
kernel name carries a 64-bit FNV-1a key of exactly those arrays, computed the
same way by the library (``wmpc_osqp_pattern_key``), so a code object only
attaches to the plan it was generated for.  Code objects are compiled with
hipcc ahead of use (``build_admm_kernel``) into ``warpmpc_amd/kernels``;
nothing is compiled while a solve runs.
"""

from __future__ import annotations

import hashlib
import os
import pathlib
import subprocess

import numpy as np

from .build import ARCH, CSRC, hipcc

PKG = pathlib.Path(__file__).resolve().parent
KERNEL_DIR = PKG / "kernels"
GEN_VERSION = 1

_FNV_OFFSET = 0xCBF29CE484222325
_FNV_PRIME = 0x100000001B3
_MASK = (1 << 64) - 1


def fnv1a64_words(words) -> int:
    """FNV-1a (64-bit) over the little-endian bytes of int32 words."""
    h = _FNV_OFFSET
    for byte in np.ascontiguousarray(np.asarray(words, dtype="<i4")).tobytes():
        h = ((h ^ byte) * _FNV_PRIME) & _MASK
    return h


def pattern_key(n: int, m: int, p, lp, li) -> int:
    """Key of the pattern data the generated iteration depends on."""
    p, lp, li = (np.asarray(a, dtype=np.int64) for a in (p, lp, li))
    words = np.concatenate([[GEN_VERSION, n, m, p.size, li.size], p, lp, li])
    return fnv1a64_words(words)


def plan_key(plan) -> int:
    q = plan.qdldl_plan
    return pattern_key(plan.n, plan.m, q.p, q.lp, q.li)


def kernel_name(key: int) -> str:
    return f"wmpc_admm_shared_{key:016x}"


# LDS of one 64-lane workgroup for KKT positions kept in LDS: 160 floats per
# lane (40 KiB) lets four workgroups, one wavefront per SIMD, share a CU.
MAX_LDS_POSITIONS = 160
# KKT positions held in registers (of the 512 a lane has at one wavefront per
# SIMD); the rest of the registers hold the substitution band, the prefetched
# operands of the entry updates and temporaries.
REGISTER_POSITIONS = 320


def default_lds_positions(n_kkt: int) -> int:
    return int(min(MAX_LDS_POSITIONS, max(0, n_kkt - REGISTER_POSITIONS)))


def admm_source(plan, *, lds_positions: int | None = None, prefetch_group: int = 2) -> tuple[str, str]:
    """(kernel name, HIP source) of the straight-line shared-factor ADMM
    kernel of an OSQP plan.

    lds_positions: KKT positions 0 .. lds_positions-1 keep their value (right-
    hand side / stash) in LDS between the sweeps instead of in registers, which
    leaves registers for prefetching; prefetch_group: backward-sweep entries
    whose update operands (x, q or z, y, l, u) are loaded one group ahead
    (2: on the mass-chain patterns at N = 10 and 20, groups of 1-4 run within
    2% of each other and 12% / 6% faster than groups of 8; larger groups hold
    more registers and slow the kernel further)."""
    q = plan.qdldl_plan
    n, m = int(plan.n), int(plan.m)
    nk = int(q.n)
    p = np.asarray(q.p, dtype=np.int64)
    lp = np.asarray(q.lp, dtype=np.int64)
    li = np.asarray(q.li, dtype=np.int64)
    if nk != n + m:
        raise ValueError("KKT dimension must be n + m")
    nl = default_lds_positions(nk) if lds_positions is None else int(lds_positions)
    if not 0 <= nl <= min(nk, MAX_LDS_POSITIONS):
        raise ValueError(f"lds_positions must be in [0, {min(nk, MAX_LDS_POSITIONS)}]")
    if prefetch_group < 1:
        raise ValueError("prefetch_group must be positive")
    name = kernel_name(plan_key(plan))
    # position c -> ('v', original variable, vtab row) or ('c', constraint, ctab row)
    kinds = []
    nv = nc = 0
    for c in range(nk):
        o = int(p[c])
        if o < n:
            kinds.append(("v", o, nv))
            nv += 1
        else:
            kinds.append(("c", o - n, nc))
            nc += 1
    resident = np.zeros(nk, dtype=bool)
    resident[:nl] = True
    # backward values needed by a later (smaller) column's gather
    needed = np.zeros(nk, dtype=bool)
    needed[li] = True

    out = []
    emit = out.append
    emit("// Generated by warpmpc_amd/codegen.py: straight-line ADMM iteration of one")
    emit(f"// KKT pattern (n={n}, m={m}, nnz_L={li.size}, {nl} positions in LDS); do not edit.")
    emit('#include "wmpc_osqp_device.h"')
    emit("namespace wmpc {")
    emit("namespace {")
    emit("__device__ __forceinline__ int64_t opaque_zero() {")
    emit("  int64_t z;")
    emit('  asm volatile("s_mov_b64 %0, 0" : "=s"(z));')
    emit("  return z;")
    emit("}")
    emit("}  // namespace")
    emit("}  // namespace wmpc")
    emit("using namespace wmpc;")
    emit(f'extern "C" __global__ __launch_bounds__(kWave) void {name}(')
    emit("    OsqpDev dv, const float* __restrict__ pv, const float* __restrict__ av, int64_t pm_ld,")
    emit("    const float* __restrict__ q, const float* __restrict__ l, const float* __restrict__ u,")
    emit("    const float* __restrict__ x0, const float* __restrict__ z0, const float* __restrict__ y0,")
    emit("    const float* __restrict__ lx, const float* __restrict__ dinv, float* __restrict__ x,")
    emit("    float* __restrict__ z, float* __restrict__ y, float* __restrict__ prim,")
    emit("    float* __restrict__ dual, float* __restrict__ obj, int32_t max_iter, int64_t batch,")
    emit("    int64_t ld, int32_t ipw, const float* __restrict__ vtab, const float* __restrict__ ctab) {")
    if nl:
        emit(f"  __shared__ float kv[{nl} * kWave];  // [position][lane]")
        emit("  float* const kvl = kv + threadIdx.x;")
    emit("  int64_t b0;")
    emit("  const int64_t b = lane_instance(ipw, batch, &b0);")
    emit("  admm_initial(dv, x0, z0, y0, Strided{x, ld, b}, Strided{z, ld, b}, Strided{y, ld, b}, ld, b);")
    emit("  const float alpha = dv.alpha, oma = 1.0f - dv.alpha, sigma = dv.sigma;")
    emit(f"  float {', '.join(f'r{c}' for c in range(nk))};")

    def tables(indent: str, zero: str):
        # wave-uniform pointers and the entry stride offset by an opaque zero:
        # loads through them stay inside the iteration instead of being hoisted
        # out of the loop, and the row offsets o * ld are formed where they are
        # used instead of being kept in scalar registers across the loop
        emit(f"{indent}const int64_t {zero} = opaque_zero();")
        emit(f"{indent}const int64_t ldi = ld + {zero};")
        # (the update tables come in as restrict arguments: through the plan
        # struct the compiler could not prove them unwritten and would read
        # them with vector loads instead of scalar ones)
        emit(f"{indent}const float* __restrict__ VT = vtab + {zero};")
        emit(f"{indent}const float* __restrict__ CT = ctab + {zero};")
        emit(f"{indent}const CStrided Q{{q, ldi, b}}, Lo{{l, ldi, b}}, Up{{u, ldi, b}};")
        emit(f"{indent}const Strided X{{x, ldi, b}}, Z{{z, ldi, b}}, Y{{y, ldi, b}};")
        emit(f"{indent}const float* __restrict__ L = lx + {zero};")
        if nl:
            emit(f"{indent}float* const KV = kvl + {zero};")

    def rhs_store(indent: str, c: int, expr: str):
        emit(f"{indent}{'KV[' + str(c) + ' * kWave]' if resident[c] else f'r{c}'} = {expr};")

    # right-hand side of the scaled initial iterate (admm_pass<true, false>)
    emit("  {")
    tables("    ", "o0")
    for c, (kind, o, t) in enumerate(kinds):
        if kind == "v":
            rhs_store("    ", c, f"sigma * X({o}) - VT[{4 * t + 3}] * Q({o})")
        else:
            rhs_store("    ", c, f"Z({o}) - CT[{8 * t + 3}] * Y({o})")
    emit("  }")
    emit("  for (int32_t it = 0; it < max_iter; ++it) {")
    emit("  {")
    tables("    ", "o0")
    emit("    const float* __restrict__ D = dinv + o0;")
    emit("    // forward substitution (column form); r_c becomes the D^-1-scaled result")
    loaded = np.zeros(nk, dtype=bool)

    def touch(i: int):
        if resident[i] and not loaded[i]:
            emit(f"    r{i} = KV[{i} * kWave];")
            loaded[i] = True

    for c in range(nk):
        touch(c)
        lo_, hi_ = int(lp[c]), int(lp[c + 1])
        for k in range(lo_, hi_):
            touch(int(li[k]))
        if hi_ > lo_:
            emit(f"    {{ const float t = r{c};")
            for k in range(lo_, hi_):
                emit(f"      r{int(li[k])} -= L[{k}] * t;")
            emit(f"      {'KV[' + str(c) + ' * kWave]' if resident[c] else f'r{c}'} = t * D[{c}]; }}")
        else:
            emit(f"    {'KV[' + str(c) + ' * kWave]' if resident[c] else f'r{c}'} = r{c} * D[{c}];")
    emit("  }")
    emit("  {")
    emit("    // backward substitution (column gather form) fused with the entry updates;")
    emit("    // its own opaque base keeps the forward sweep's L loads (and the stash")
    emit("    // stores to LDS) from being reused here, which would hold them in")
    emit("    // registers between the sweeps")
    tables("    ", "o1")
    wdecl = [f"w{c}" for c in range(nk) if needed[c]]
    if wdecl:
        emit(f"    float {', '.join(wdecl)};")
    order = list(range(nk - 1, -1, -1))
    groups = [order[i:i + prefetch_group] for i in range(0, nk, prefetch_group)]

    def loads(group):
        # update operands of a group, issued one group ahead of their use
        for c in group:
            kind, o, _ = kinds[c]
            if kind == "v":
                emit(f"    const float px{c} = X({o}), pq{c} = Q({o});")
            else:
                emit(f"    const float pz{c} = Z({o}), py{c} = Y({o}), pl{c} = Lo({o}), pu{c} = Up({o});")
            if resident[c]:
                emit(f"    const float ps{c} = KV[{c} * kWave];")

    loads(groups[0])
    for j, group in enumerate(groups):
        if j + 1 < len(groups):
            loads(groups[j + 1])
        for c in group:
            lo_, hi_ = int(lp[c]), int(lp[c + 1])
            w = f"w{c}" if needed[c] else "wc"
            lhs = w if needed[c] else "const float wc"
            sv = f"ps{c}" if resident[c] else f"r{c}"
            if hi_ > lo_:
                terms = [f"L[{k}] * w{int(li[k])}" for k in range(lo_, hi_)]
                emit(f"    {{ float a = {terms[0]};")
                for tm in terms[1:]:
                    emit(f"      a += {tm};")
                emit(f"      {lhs} = {sv} - a;")
            else:
                emit(f"    {{ {lhs} = {sv};")
            kind, o, t = kinds[c]
            if kind == "v":
                emit(f"      const float xn = alpha * {w} + oma * px{c}; X({o}) = xn;")
                rhs_store("      ", c, f"sigma * xn - VT[{4 * t + 3}] * pq{c}")
            else:
                emit(f"      const float rinv = CT[{8 * t + 3}], rho = CT[{8 * t + 4}], e = CT[{8 * t + 5}];")
                emit(f"      const float rz = pz{c} - rinv * py{c};")
                emit(f"      const float zt = rz + rinv * {w};")
                emit(f"      const float zr = alpha * zt + oma * pz{c};")
                emit(f"      const float zn = nan_min(nan_max(zr + rinv * py{c}, pl{c} * e), pu{c} * e);")
                emit(f"      const float yn = py{c} + rho * (zr - zn);")
                emit(f"      Z({o}) = zn; Y({o}) = yn;")
                rhs_store("      ", c, "zn - rinv * yn")
            emit("    }")
    emit("  }")
    emit("  }")
    emit("  admm_finish(dv, pv, av, pm_ld, CStrided{q, ld, b}, Strided{x, ld, b}, Strided{z, ld, b},")
    emit("              Strided{y, ld, b}, prim, dual, obj, b);")
    emit("}")
    src = "\n".join(out) + "\n"
    return name, src


def code_object_path(plan, directory: pathlib.Path | None = None) -> pathlib.Path:
    return (directory or KERNEL_DIR) / f"{kernel_name(plan_key(plan))}.hsaco"


def build_admm_kernel(plan, *, directory: pathlib.Path | None = None, force: bool = False,
                      verbose: bool = False) -> pathlib.Path:
    """Generates and compiles the plan's straight-line kernel (hipcc, gfx950)
    into a code object; returns its path (kept when it already exists)."""
    from .build import HIPCC_FLAGS

    path = code_object_path(plan, directory)
    if path.exists() and not force:
        return path
    path.parent.mkdir(parents=True, exist_ok=True)
    name, src = admm_source(plan)
    digest = hashlib.sha1(src.encode()).hexdigest()[:12]
    hip = path.with_name(f"{name}.{digest}.hip")
    hip.write_text(src)
    tmp = path.with_suffix(f".hsaco.tmp{os.getpid()}")
    flags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC", "-Wall")]
    cmd = [hipcc(), *flags, "--genco", f"--offload-arch={ARCH}", f"-I{CSRC}", "-o", str(tmp), str(hip)]
    if verbose:
        print(" ".join(cmd), flush=True)
    try:
        subprocess.run(cmd, check=True)
        os.replace(tmp, path)
    finally:
        hip.unlink(missing_ok=True)
        tmp.unlink(missing_ok=True)
    return path


def find_admm_kernel(plan, directory: pathlib.Path | None = None) -> pathlib.Path | None:
    path = code_object_path(plan, directory)
    return path if path.exists() else None


__all__ = [
    "GEN_VERSION",
    "KERNEL_DIR",
    "admm_source",
    "build_admm_kernel",
    "code_object_path",
    "find_admm_kernel",
    "fnv1a64_words",
    "kernel_name",
    "pattern_key",
    "plan_key",
]
