"""Per-pattern straight-line HIP code for the batched LDLᵀ (gfx950).

For one fixed pattern the whole up-looking factorization is known at plan time:
every row's reach, every update ``y_r -= L[r, c] y_c``, every value slot.  The
generator emits it as straight-line code in which

* every L entry, reciprocal pivot and forward-substitution value that a later
  row still needs is a named local (register-allocated by the compiler); with
  the largest-subtree-first postorder of the planner these live sets are
  bounded by the band of the matrix (a few hundred values), independent of its
  size;
* all indices are immediates; global accesses are one uniform (scalar) base
  per entry plus the lane offset, so every access of a wavefront is one
  contiguous 256-byte segment of the batch-major arrays;
* nothing but the compulsory streams touches HBM: values and rhs in, L / D out,
  L and the scaled forward result re-read once by the backward sweep, x out.

Kernels (all one lane per instance, batch-major [entry][ld] arrays):
  ``factor``        a -> lx, d, dinv                     (CompiledQDLDL.factor)
  ``factor_solve``  a, rhs -> x, lx, d  (+ work [n][ld])  (factor_and_solve)

The row program follows ``warpmpc/jax_qdldl/core.py:867-933``; the fused
forward substitution is the row-oriented form ``w_k = b_k - sum_c L[k,c] w_c``.
"""

from __future__ import annotations

import hashlib
import io

import numpy as np

from .qdldl.tables import ragged_program
from .qdldl.types import QDLDLPlan

CODEGEN_VERSION = 1


def pattern_key(plan: QDLDLPlan) -> str:
    """Stable hash of everything the generated code depends on."""
    h = hashlib.sha256()
    h.update(f"v{CODEGEN_VERSION}".encode())
    for name, arr in ragged_program(plan).items():
        h.update(name.encode())
        h.update(np.ascontiguousarray(arr, dtype=np.int32).tobytes())
    return h.hexdigest()[:16]


def _lfmt(v: float) -> str:
    return f"{v!r}f"


class _Emitter:
    def __init__(self):
        self.buf = io.StringIO()

    def __call__(self, line: str = "") -> None:
        self.buf.write(line)
        self.buf.write("\n")

    def text(self) -> str:
        return self.buf.getvalue()


def _row_programs(plan: QDLDLPlan):
    t = ragged_program(plan)
    n = plan.n
    lp, li = t["lp"], t["li"]
    rows = []
    slot_of = np.full(n, -1, dtype=np.int64)
    for k in range(n):
        r0, r1 = int(t["proc_ptr"][k]), int(t["proc_ptr"][k + 1])
        cols = t["proc_cols"][r0:r1].tolist()
        for s, c in enumerate(cols):
            slot_of[c] = s
        init = [(int(slot_of[t["init_rows"][j]]), int(t["perm_to_orig"][t["init_pos"][j]]))
                for j in range(int(t["init_ptr"][k]), int(t["init_ptr"][k + 1]))]
        procs = []
        for s, c in enumerate(cols):
            off = int(t["proc_off"][r0 + s])
            upd = [(int(slot_of[li[lp[c] + q]]), int(lp[c] + q)) for q in range(off)]
            procs.append((c, int(t["proc_lidx"][r0 + s]), upd))
        diag = int(t["perm_to_orig"][t["diag_pos"][k]])
        rows.append((diag, init, procs))
        for c in cols:
            slot_of[c] = -1
    return t, rows


def _last_use(plan: QDLDLPlan, t) -> tuple[np.ndarray, np.ndarray]:
    """Last factor row that reads column c's (dinv, w) and L entry l."""
    n = plan.n
    lp, li = t["lp"], t["li"]
    col_last = np.full(n, -1, dtype=np.int64)
    for c in range(n):
        if lp[c + 1] > lp[c]:
            col_last[c] = li[lp[c + 1] - 1]
    l_last = np.full(plan.nnz_l, -1, dtype=np.int64)
    for c in range(n):
        for q in range(int(lp[c]), int(lp[c + 1]) - 1):
            l_last[q] = col_last[c]
    return col_last, l_last


def generate(plan: QDLDLPlan, kinds=("factor", "factor_solve")) -> str:
    """HIP source with one ``extern "C"`` kernel per requested kind."""
    key = pattern_key(plan)
    t, rows = _row_programs(plan)
    n = plan.n
    e = _Emitter()
    e(f"// generated by warpmpc_amd.codegen v{CODEGEN_VERSION} for pattern {key}")
    e(f"// n = {n}, nnz_a = {plan.nnz_a}, nnz_l = {plan.nnz_l}")
    e("#include <hip/hip_runtime.h>")
    e("#include <stdint.h>")
    # uniform (scalar) base of entry i for this workgroup's 64 instances, plus the
    # 32-bit lane offset: selects the scalar-base global addressing form
    e("#define LD(p, i) ((p) + (int64_t)(i) * ld)[tid]")
    e("#define ST(p, i, v) (((p) + (int64_t)(i) * ld)[tid] = (v))")
    for kind in kinds:
        if kind == "factor":
            _emit_factor(e, plan, t, rows, key, fuse_solve=False)
        elif kind == "factor_solve":
            _emit_factor(e, plan, t, rows, key, fuse_solve=True)
        else:
            raise ValueError(f"unknown kernel kind {kind!r}")
    return e.text()


def kernel_name(kind: str, key: str) -> str:
    return f"wmpc_gen_{kind}_{key}"


def _emit_factor(e: _Emitter, plan: QDLDLPlan, t, rows, key: str, *, fuse_solve: bool) -> None:
    n = plan.n
    lp, li = t["lp"], t["li"]
    p = t["p"]
    kind = "factor_solve" if fuse_solve else "factor"
    if fuse_solve:
        # outputs are deliberately not __restrict__: the backward sweep must re-load L
        # and the scaled forward result from memory instead of keeping them live
        sig = ("const float* __restrict__ a, const float* __restrict__ rhs, float* x, "
               "float* lx, float* d, float* work, int64_t batch, int64_t ld")
    else:
        sig = ("const float* __restrict__ a, float* lx, float* d, float* dinv, "
               "int64_t batch, int64_t ld")
    e(f'extern "C" __global__ __launch_bounds__(64, 1) void {kernel_name(kind, key)}({sig}) {{')
    e("  const unsigned tid = threadIdx.x;")
    e("  const int64_t b0 = (int64_t)blockIdx.x * 64;")
    e("  if (b0 + tid >= batch) return;")
    e("  a += b0; lx += b0; d += b0;")
    if fuse_solve:
        e("  rhs += b0; x += b0; work += b0;")
    else:
        e("  dinv += b0;")
    for k, (diag, init, procs) in enumerate(rows):
        e(f"  // row {k}")
        e("  __builtin_amdgcn_sched_barrier(0);")
        e("  {")
        touched = {}
        for slot, src in init:
            if slot in touched:
                e(f"    y{slot} += LD(a, {src});")
            else:
                e(f"    float y{slot} = LD(a, {src});")
                touched[slot] = True
        for s, (c, dest, upd) in enumerate(procs):
            for tgt, lidx in upd:
                if tgt not in touched:
                    e(f"    float y{tgt} = 0.0f;")
                    touched[tgt] = True
        e(f"    float dk = LD(a, {diag});")
        if fuse_solve:
            e(f"    float fw = LD(rhs, {int(p[k])});")
        for s, (c, dest, upd) in enumerate(procs):
            val = f"y{s}" if s in touched else "0.0f"
            e(f"    {{ const float v = {val};")
            for tgt, lidx in upd:
                e(f"      y{tgt} -= L{lidx} * v;")
            e(f"      L{dest} = v * DI{c}; ST(lx, {dest}, L{dest}); dk -= v * L{dest};")
            if fuse_solve:
                e(f"      fw -= L{dest} * W{c};")
            e("    }")
        e(f"    DI{k} = 1.0f / dk; ST(d, {k}, dk);")
        if fuse_solve:
            e(f"    W{k} = fw; ST(work, {k}, fw * DI{k});")
        else:
            e(f"    ST(dinv, {k}, DI{k});")
        e("  }")
    if fuse_solve:
        # backward sweep: z_c = w_c * dinv_c (stored), x_c = z_c - sum_r L[r, c] x_r
        e("  __builtin_amdgcn_fence(__ATOMIC_RELEASE, \"workgroup\");")
        # opaque copies: backward addresses must not be CSE'd with the forward ones
        e("  const float* lxb = lx; const float* wkb = work;")
        e("  asm volatile(\"\" : \"+s\"(lxb), \"+s\"(wkb) : : \"memory\");")
        for c in range(n - 1, -1, -1):
            q0, q1 = int(lp[c]), int(lp[c + 1])
            if q1 > q0:
                terms = " + ".join(f"LD(lxb, {q}) * X{int(li[q])}" for q in range(q0, q1))
                e(f"  const float X{c} = LD(wkb, {c}) - ({terms});")
            else:
                e(f"  const float X{c} = LD(wkb, {c});")
            e(f"  ST(x, {int(p[c])}, X{c});")
    e("}")
    # declare the function-scope locals: prepend after signature (C++ needs them before use)
    txt = e.buf.getvalue()
    head, sep, body = txt.rpartition(f"void {kernel_name(kind, key)}(")
    decl_lines = []
    decl_lines.append("  float " + ", ".join(f"L{i}" for i in range(plan.nnz_l)) + ";" if plan.nnz_l else "")
    decl_lines.append("  float " + ", ".join(f"DI{i}" for i in range(n)) + ";")
    if fuse_solve:
        decl_lines.append("  float " + ", ".join(f"W{i}" for i in range(n)) + ";")
    first_brace = body.index("{\n") + 2
    body = body[:first_brace] + "\n".join(x for x in decl_lines if x) + "\n" + body[first_brace:]
    e.buf = io.StringIO()
    e.buf.write(head + sep + body)


__all__ = ["CODEGEN_VERSION", "generate", "kernel_name", "pattern_key"]
