"""CPU replay of the HIP kernels' row/column programs (test helper, no GPU).

The HIP library derives, per pattern, programs whose operands are LDS slots
(reach slots Y, column window CW/FW, live-L slots AL, backward window BW) or L
entries re-read from global memory.  This module replays those
programs for one instance in float64 with one Python dict per LDS region, so a
slot that is reused while still live, or read before it is written, changes
the result and fails the comparison with the oracle.  ``replay_factor_solve``
mirrors ``factor_rows`` with the fused forward substitution
(``wmpc_qdldl_device.h``); ``replay_record_sweeps`` mirrors the chunked
``fwd_cols`` / ``bwd_cols`` of ``wmpc_stream.h``.
"""

from __future__ import annotations

import ctypes

import numpy as np

from warpmpc_amd import _lib
from warpmpc_amd.qdldl.tables import ragged_program


class HostPlan:
    """A host-only ``wmpc_qdldl_plan`` and its program tables."""

    def __init__(self, plan, lds_budget: int | None = None):
        self.handle = None
        lib = _lib.load()
        self.plan = plan
        self.tables = ragged_program(plan)
        pat = _lib.QdldlPattern()
        pat.n, pat.nnz_a, pat.nnz_l = plan.n, plan.nnz_a, plan.nnz_l
        for name, arr in self.tables.items():
            setattr(pat, name, _lib.i32p(arr))
        self._pat = pat
        h = ctypes.c_void_p()
        _lib.check(lib.wmpc_qdldl_plan_create_host(ctypes.byref(pat), ctypes.byref(h)), "plan_create_host")
        self.handle = h
        if lds_budget is not None:
            _lib.check(lib.wmpc_qdldl_plan_set_lds_budget(h, int(lds_budget)), "set_lds_budget")
        self.slots = _lib.plan_slots(h)
        self.row = _lib.plan_program(h, "row").reshape(-1, 4)
        self.init = _lib.plan_program(h, "init").reshape(-1, 2)
        self.proc = _lib.plan_program(h, "proc").reshape(-1, 4)
        self.upd = _lib.plan_program(h, "upd").reshape(-1, 2)
        self.bcol = _lib.plan_program(h, "bcol").reshape(-1, 2)
        self.bent = _lib.plan_program(h, "bent")
        self.fcol = _lib.plan_program(h, "fcol").reshape(-1, 4)
        self.bcol2 = _lib.plan_program(h, "bcol2").reshape(-1, 4)
        self.fent = _lib.plan_program(h, "fent")
        self.ftpos = _lib.plan_program(h, "ftpos")
        self.frec = _lib.plan_program(h, "frec").reshape(-1, 64, 2)
        self.brec = _lib.plan_program(h, "brec").reshape(-1, 64, 2)

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            _lib.load().wmpc_qdldl_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class _Region:
    """One LDS region: slot -> value, with a bound check and a write check."""

    def __init__(self, name: str, size: int):
        self.name, self.size, self.v = name, size, {}

    def __getitem__(self, s):
        if not 0 <= s < self.size:
            raise IndexError(f"{self.name} slot {s} outside [0, {self.size})")
        if s not in self.v:
            raise KeyError(f"{self.name} slot {s} read before any write")
        return self.v[s]

    def __setitem__(self, s, val):
        if not 0 <= s < self.size:
            raise IndexError(f"{self.name} slot {s} outside [0, {self.size})")
        self.v[s] = val


FIRST_TOUCH = 1 << 16
MARK = 1 << 20
NO_SLOT = 0xFFFF
V_PER_CHUNK = 16
L_PER_CHUNK = 64


def replay_record_sweeps(hp: HostPlan, lx, dinv, rhs_orig):
    """``fwd_cols`` + ``bwd_cols`` over the chunked record streams.

    Each chunk's operands are taken from windows [base, base + 16) (values)
    and [base, base + 64) (L) exactly as the kernel stages them, so a record
    that consumes a value outside its chunk's staged window fails here.
    """
    n, nl = hp.plan.n, hp.plan.nnz_l
    p = hp.tables["p"]
    R = np.zeros(n)
    R[hp.ftpos] = rhs_orig[p]

    def bases(recs, first):
        nchunks = recs.shape[0]
        out = [(0, 0, 0), tuple(first)]
        for t in range(nchunks - 2):
            h0, h1 = recs[t][0], recs[t][1]
            out.append((int(h0[0]), int(h0[1]), int(h1[0])))
        return out

    # first-chunk offsets are not exported: recompute from the record counts
    def first_offsets(recs, lfn, vfn, dfn):
        r = recs[0]
        nrec = int(r[1][1])
        l = v = d = 0
        for k in range(2, 2 + nrec):
            l += lfn(r[k]); v += vfn(r[k]); d += dfn(r[k])
        return (l, v, d)

    is_mark = lambda r: bool(int(r[0]) & MARK)
    f_first = first_offsets(hp.frec, lambda r: 0 if is_mark(r) else 1,
                            lambda r: (int(int(r[0]) & 0xFFFF) == NO_SLOT) if is_mark(r) else (int(r[0]) >> 16) & 1,
                            lambda r: 1 if is_mark(r) else 0)
    W = _Region("W", hp.slots["backwin"])
    stash = np.zeros(n)
    li = vi = di = col = 0
    val = 0.0
    fb = bases(hp.frec, f_first)
    for t in range(hp.frec.shape[0]):
        lb, vb, db = fb[t]
        chunk = hp.frec[t]
        nrec = int(chunk[1][1])
        for k in range(2, 2 + nrec):
            w = int(chunk[k][0])
            if w & MARK:
                slot = w & 0xFFFF
                if slot != NO_SLOT:
                    val = W[slot]
                else:
                    assert 0 <= vi - vb < V_PER_CHUNK
                    val = R[vi]
                    vi += 1
                assert 0 <= di - db < V_PER_CHUNK
                stash[col] = val * dinv[di]
                di += 1
                col += 1
            else:
                assert 0 <= li - lb < L_PER_CHUNK
                slot = w & 0xFFFF
                if w & FIRST_TOUCH:
                    assert 0 <= vi - vb < V_PER_CHUNK
                    W[slot] = R[vi] - lx[li] * val
                    vi += 1
                else:
                    W[slot] = W[slot] - lx[li] * val
                li += 1
    assert (li, vi, di, col) == (nl, n, n, n)
    b_first = first_offsets(hp.brec, lambda r: 0 if is_mark(r) else 1, lambda r: 1 if is_mark(r) else 0,
                            lambda r: 0)
    W = _Region("W", hp.slots["backwin"])
    x = np.zeros(n)
    li = vi = 0
    col = n - 1
    corr = 0.0
    bb = bases(hp.brec, b_first)
    for t in range(hp.brec.shape[0]):
        lb, vb, _ = bb[t]
        chunk = hp.brec[t]
        nrec = int(chunk[1][1])
        for k in range(2, 2 + nrec):
            w = int(chunk[k][0])
            if w & MARK:
                assert 0 <= vi - vb < V_PER_CHUNK
                wv = stash[n - 1 - vi] - corr
                vi += 1
                corr = 0.0
                slot = w & 0xFFFF
                if slot != NO_SLOT:
                    W[slot] = wv
                assert int(chunk[k][1]) == p[col]
                x[p[col]] = wv
                col -= 1
            else:
                assert 0 <= li - lb < L_PER_CHUNK
                corr += lx[nl - 1 - li] * W[w & 0xFFFF]
                li += 1
    assert (li, vi, col) == (nl, n, -1)
    return x


def replay_factor_solve(hp: HostPlan, a: np.ndarray, rhs: np.ndarray):
    """factor_rows<FWD=true> + backward_cols for one instance.

    a: [nnz_a] values in original upper-CSC order; rhs: [n] in original order.
    Returns (x, lx, d, dinv) in float64.
    """
    n, nl = hp.plan.n, hp.plan.nnz_l
    p = hp.tables["p"]
    s = hp.slots
    Y = _Region("Y", s["reach"])
    CW = _Region("CW", s["colwin"])
    FW = _Region("FW", s["colwin"])
    AL = _Region("AL", s["live_l"])
    lx = np.full(nl, np.nan)
    d = np.zeros(n)
    dinv = np.zeros(n)
    stash = np.zeros(n)
    r = j = u = 0
    for k in range(n):
        r1, i1, ck, dsrc = (int(v) for v in hp.row[k])
        for t in range(r1 - r):
            Y[t] = 0.0
        while j < i1:
            ys, src = hp.init[j]
            Y[int(ys)] = Y[int(ys)] + a[src]
            j += 1
        dk = a[dsrc]
        fw = rhs[p[k]]
        t = 0
        while r < r1:
            packed, lidx, split, uend = (int(v) for v in hp.proc[r])
            cs, keep = packed & 0xFFFF, (packed >> 16) - 1
            if not u <= split <= uend:
                raise AssertionError(f"reach entry {r}: global split {split} outside [{u}, {uend}]")
            val = Y[t]
            while u < uend:
                ys, ls = (int(v) for v in hp.upd[u])
                if (ls >= 0) != (u < split):
                    raise AssertionError(f"update {u}: L source on the wrong side of the split {split}")
                lv = AL[ls] if ls >= 0 else lx[-ls - 1]
                if np.isnan(lv):
                    raise AssertionError(f"update {u} reads L slot {-ls - 1} before it is written")
                Y[ys] = Y[ys] - lv * val
                u += 1
            lval = val * CW[cs]
            lx[lidx] = lval
            if keep >= 0:
                AL[keep] = lval
            dk -= val * lval
            fw -= lval * FW[cs]
            r += 1
            t += 1
        d[k] = dk
        dinv[k] = 1.0 / dk
        if ck >= 0:
            CW[ck] = dinv[k]
            FW[ck] = fw
        stash[k] = fw
    x = replay_backward(hp, lx, dinv, stash)
    return x, lx, d, dinv


def replay_backward(hp: HostPlan, lx, dinv, stash):
    n = hp.plan.n
    p = hp.tables["p"]
    BW = _Region("BW", hp.slots["backwin"])
    x = np.zeros(n)
    for c in range(n - 1, -1, -1):
        bs, qend = (int(v) for v in hp.bcol[c])
        q = int(hp.bcol[c - 1][1]) if c > 0 else 0
        acc = stash[c] * dinv[c]
        for qq in range(q, qend):
            acc -= lx[qq] * BW[int(hp.bent[qq])]
        if bs >= 0:
            BW[bs] = acc
        x[p[c]] = acc
    return x
