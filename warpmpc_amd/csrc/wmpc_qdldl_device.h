// Device-side building blocks of the batched fixed-pattern LDLᵀ (gfx950).
//
// One lane owns one instance; one wavefront (64 lanes) is one workgroup.  Every
// per-instance array in global memory is batch-major ([entry][ld]), so a
// wavefront touching entry i of 64 consecutive instances issues one contiguous
// 256-byte access.  Program tables are wave-uniform: the factorization reads
// them as streams of 64-entry chunks (one coalesced load a chunk ahead, an
// entry per v_readlane), so it needs whole wavefronts.  Working
// values live in LDS slots laid out [slot][64]: lane t of slot s is
// lds[s * 64 + t], so one slot access of a wavefront hits 64 distinct banks.
//
// The factorization is the up-looking QDLDL recurrence of
// warpmpc/jax_qdldl/core.py:867-933 (== warp_backend.py:41-173); the solves are
// x = Pᵀ L⁻ᵀ D⁻¹ L⁻¹ P b (core.py:935-1028) with the forward substitution in
// row (gather) form and the backward substitution in column (gather) form.
#pragma once

#include "wmpc_common.h"

namespace wmpc {

constexpr int kWave = 64;

// Instance of this lane.  A wavefront serves `ipw` instances (a power of two,
// 8..64): lane l works on instance blockIdx.x * ipw + (l mod ipw), so with
// ipw < 64 the 64 / ipw lanes of one instance do identical work (identical
// loads, identical values stored to identical addresses).  Fewer instances per
// wavefront give a small batch more wavefronts, one per SIMD.  Lanes past the
// batch take the last instance.  *b0 receives the unclamped index.
__device__ __forceinline__ int64_t lane_instance(int32_t ipw, int64_t batch, int64_t* b0) {
  const int64_t i = (int64_t)blockIdx.x * ipw + (int64_t)(threadIdx.x & (ipw - 1));
  *b0 = i;
  return i < batch ? i : batch - 1;
}

// Device view of one pattern's programs (see wmpc_program.h for their meaning).
struct QdldlProg {
  int32_t n = 0, nnz_a = 0, nnz_l = 0;
  int32_t ymax = 0, ncol = 0, nact = 0, nbw = 0;
  const int4* row = nullptr;   // [n] {proc_end, init_end, colslot_k, diag_src}
  const int2* init = nullptr;  // {yslot, src}
  const int4* proc = nullptr;  // {colslot(c) | (AL slot kept + 1) << 16, lidx, first global update, upd_end}
  const int2* upd = nullptr;   // {yslot, AL slot or -(lidx+1)}
  const int2* bcol = nullptr;  // [n] {bslot_c or -1, lp[c+1]}
  const int32_t* bent = nullptr;  // [nnz_l]
  const int32_t* p = nullptr;     // [n]
  // column sweeps (wmpc_stream.h)
  const int4* fcol = nullptr;     // [n]
  const int4* bcol2 = nullptr;    // [n]
  const int32_t* fent = nullptr;  // [nnz_l]
  const int32_t* rperm = nullptr; // [n] original index of the i-th first-touch row
  // chunked record streams of the sweeps and the stream offsets of chunk 1
  const int32_t* frec = nullptr;
  const int32_t* brec = nullptr;
  int32_t fchunks = 0, bchunks = 0;
  int32_t f1[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
  // LDS columns of the slot windows: 64 (one per lane), or 32/16/8 for patterns
  // whose windows do not fit otherwise; lane l uses column l mod lcols, so a
  // launch serves at most lcols instances per wavefront
  int32_t lcols = kWave;
  // instances per wavefront chosen for this plan (0: the process-wide choice)
  int32_t ipw = 0;
  // entries of the init / proc / upd tables
  int32_t ninit = 0, nproc = 0, nupd = 0;
};

__device__ __forceinline__ int32_t read_lane(int32_t v, int32_t j) {
  return __builtin_amdgcn_readlane(v, j);
}
__device__ __forceinline__ float read_lane(float v, int32_t j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ int2 read_lane(int2 v, int32_t j) {
  return make_int2(__builtin_amdgcn_readlane(v.x, j), __builtin_amdgcn_readlane(v.y, j));
}
__device__ __forceinline__ int4 read_lane(int4 v, int32_t j) {
  return make_int4(__builtin_amdgcn_readlane(v.x, j), __builtin_amdgcn_readlane(v.y, j),
                   __builtin_amdgcn_readlane(v.z, j), __builtin_amdgcn_readlane(v.w, j));
}

// Sequential reader of a wave-uniform table of `count` entries: lane l holds
// entry base + l of the current 64-entry chunk and base + 64 + l of the next,
// so one coalesced load, issued a chunk ahead, serves 64 reads, and an entry
// reaches scalar registers by v_readlane instead of a load per read.  Reads
// must come in nondecreasing order, and every lane of the wavefront must be
// active (v_readlane reads lanes regardless of the execution mask).
template <class T>
struct TableStream {
  const T* t;
  int32_t count, base;
  T cur, nxt;
  __device__ __forceinline__ T fetch(int32_t e) const {
    return e < count ? t[e] : T{};
  }
  __device__ __forceinline__ TableStream(const T* table, int32_t n, int32_t first)
      : t(table), count(n), base(first) {
    const int32_t l = (int32_t)(threadIdx.x & (kWave - 1));
    cur = fetch(first + l);
    nxt = fetch(first + kWave + l);
  }
  __device__ __forceinline__ T operator[](int32_t i) {
    while (i - base >= kWave) {
      base += kWave;
      cur = nxt;
      nxt = fetch(base + kWave + (int32_t)(threadIdx.x & (kWave - 1)));
    }
    return read_lane(cur, i - base);
  }
  // Entries i .. i+3 (one range check when they lie in the current chunk).
  __device__ __forceinline__ void get4(int32_t i, T& a, T& b, T& c, T& d) {
    const int32_t j = i - base;
    if (j + 3 < kWave) {
      a = read_lane(cur, j);
      b = read_lane(cur, j + 1);
      c = read_lane(cur, j + 2);
      d = read_lane(cur, j + 3);
    } else {
      a = (*this)[i];
      b = (*this)[i + 1];
      c = (*this)[i + 2];
      d = (*this)[i + 3];
    }
  }
};

// LDS slots of the factorization (fused forward adds one window) and the
// floats they take in a workgroup.
__host__ __device__ inline int32_t factor_lds_slots(const QdldlProg& P, bool fwd) {
  return P.ymax + P.ncol * (fwd ? 2 : 1) + P.nact;
}
__host__ __device__ inline int64_t factor_lds_floats(const QdldlProg& P, bool fwd) {
  return (int64_t)factor_lds_slots(P, fwd) * P.lcols;
}

// Values straight from a [nnz_a][ld] array (CompiledQDLDL.factor input).
struct DirectSource {
  const float* __restrict__ a;
  int64_t ld;
  __device__ __forceinline__ float operator()(int32_t slot, int64_t b) const {
    return (a + (int64_t)slot * ld)[(uint32_t)b];
  }
};

// KKT values assembled on the fly from P and A (solver.py:764-776):
// kind 0: p[idx] * scale + off, kind 1: a[idx] * scale, kind 2: off.
// P/A are [nnz][pm_ld] per instance, or one shared [nnz] column when pm_ld == 0.
struct KktSource {
  const float* __restrict__ p;
  const float* __restrict__ a;
  int64_t pm_ld;
  const int32_t* __restrict__ kind;
  const int32_t* __restrict__ idx;
  const float* __restrict__ scale;
  const float* __restrict__ off;
  __device__ __forceinline__ float operator()(int32_t slot, int64_t b) const {
    const int32_t k = kind[slot];
    const int64_t row = pm_ld ? (int64_t)idx[slot] * pm_ld : (int64_t)idx[slot];
    const uint32_t lane = pm_ld ? (uint32_t)b : 0u;
    if (k == 0) return (p + row)[lane] * scale[slot] + off[slot];
    if (k == 1) return (a + row)[lane] * scale[slot];
    return off[slot];
  }
};

// Strided per-instance array: element i of instance b at ptr[i * ld + b]
// (ld == 1, b == 0: one array shared by the whole batch).  The address is
// formed as a wave-uniform row base (i * ld, scalar registers) plus the lane's
// 32-bit offset b, the form the scalar-base global loads take directly.
struct Strided {
  float* ptr;
  int64_t ld;
  int64_t b;
  __device__ __forceinline__ float& operator()(int64_t i) const {
    return (ptr + i * ld)[(uint32_t)b];
  }
};
struct CStrided {
  const float* __restrict__ ptr;
  int64_t ld;
  int64_t b;
  __device__ __forceinline__ float operator()(int64_t i) const {
    return (ptr + i * ld)[(uint32_t)b];
  }
};

struct NoRhs {
  __device__ __forceinline__ float operator()(int32_t, int32_t) const { return 0.f; }
};

// Up-looking factorization of one instance (values src(slot, sb)).
// lx / d / dinv: outputs (d may have ptr == nullptr).  lds: this lane's slot
// column base.  FWD additionally runs the forward substitution
// w[k] = rhs(k) - sum_c L[k,c] w[c] with the freshly computed row of L and
// stores w into stash(k).  Called by every lane of the wavefront together
// (the program tables are read through TableStream).
template <bool FWD, class Src, class Rhs>
__device__ __forceinline__ void factor_rows(const QdldlProg& P, const Src& src, int64_t sb,
                                            const Strided& lx, const Strided& d,
                                            const Strided& dinv, float* lds, const Rhs& rhs,
                                            const Strided& stash) {
  const int32_t lc = P.lcols;
  float* Y = lds;
  float* CW = Y + P.ymax * lc;
  float* AL = CW + P.ncol * lc;
  float* FW = AL + P.nact * lc;
  TableStream<int4> rows(P.row, P.n, 0), procs(P.proc, P.nproc, 0);
  TableStream<int2> inits(P.init, P.ninit, 0), upds(P.upd, P.nupd, 0);
  int32_t r = 0, j = 0, u = 0;
  for (int32_t k = 0; k < P.n; ++k) {
    const int4 ri = rows[k];
    const int32_t r1 = ri.x;
    for (int32_t s = 0; s < r1 - r; ++s) Y[s * lc] = 0.f;
    float dk = src(ri.w, sb);
    float fw = 0.f;
    if constexpr (FWD) fw = rhs(k, P.p[k]);
    // the column's values are independent loads: issue them in groups of 8
    // before their (in-order) accumulation into the row window
    for (; j + 8 <= ri.y; j += 8) {
      int2 e[8];
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        e[i] = inits[j + i];
        v[i] = src(e[i].y, sb);
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) Y[e[i].x * lc] += v[i];
    }
    for (; j < ri.y; ++j) {
      const int2 e = inits[j];
      Y[e.x * lc] += src(e.y, sb);
    }
    for (int32_t s = 0; r < r1; ++r, ++s) {
      const int4 pe = procs[r];
      const float val = Y[s * lc];
      // the updates of one reach entry hit distinct later slots: issue their
      // loads in groups so several LDS / global round trips overlap; the
      // live-L operands (LDS) come first, then the L entries re-read from
      // global memory, so neither loop branches on the operand's source
      for (; u + 4 <= pe.z; u += 4) {
        int2 e0, e1, e2, e3;
        upds.get4(u, e0, e1, e2, e3);
        const float l0 = AL[e0.y * lc], l1 = AL[e1.y * lc], l2 = AL[e2.y * lc], l3 = AL[e3.y * lc];
        const float y0 = Y[e0.x * lc], y1 = Y[e1.x * lc];
        const float y2 = Y[e2.x * lc], y3 = Y[e3.x * lc];
        Y[e0.x * lc] = y0 - l0 * val;
        Y[e1.x * lc] = y1 - l1 * val;
        Y[e2.x * lc] = y2 - l2 * val;
        Y[e3.x * lc] = y3 - l3 * val;
      }
      for (; u < pe.z; ++u) {
        const int2 e = upds[u];
        Y[e.x * lc] -= AL[e.y * lc] * val;
      }
      for (; u + 4 <= pe.w; u += 4) {
        int2 e0, e1, e2, e3;
        upds.get4(u, e0, e1, e2, e3);
        const float l0 = lx(-(int64_t)e0.y - 1), l1 = lx(-(int64_t)e1.y - 1);
        const float l2 = lx(-(int64_t)e2.y - 1), l3 = lx(-(int64_t)e3.y - 1);
        const float y0 = Y[e0.x * lc], y1 = Y[e1.x * lc];
        const float y2 = Y[e2.x * lc], y3 = Y[e3.x * lc];
        Y[e0.x * lc] = y0 - l0 * val;
        Y[e1.x * lc] = y1 - l1 * val;
        Y[e2.x * lc] = y2 - l2 * val;
        Y[e3.x * lc] = y3 - l3 * val;
      }
      for (; u < pe.w; ++u) {
        const int2 e = upds[u];
        Y[e.x * lc] -= lx(-(int64_t)e.y - 1) * val;
      }
      const int32_t cs = pe.x & 0xFFFF, keep = (pe.x >> 16) - 1;
      const float lval = val * CW[cs * lc];
      lx(pe.y) = lval;
      if (keep >= 0) AL[keep * lc] = lval;
      dk -= val * lval;
      if constexpr (FWD) fw -= lval * FW[cs * lc];
    }
    if (d.ptr) d(k) = dk;
    const float di = 1.0f / dk;
    dinv(k) = di;
    if (ri.z >= 0) {
      CW[ri.z * lc] = di;
      if constexpr (FWD) FW[ri.z * lc] = fw;
    }
    // the backward sweep consumes the D⁻¹-scaled forward result
    if constexpr (FWD) stash(k) = fw * di;
  }
}

// rhs read from a strided array in original order.
struct ArrayRhs {
  CStrided rhs;
  __device__ __forceinline__ float operator()(int32_t, int32_t o) const { return rhs(o); }
};

// solution written to a strided array in original order.
struct ArrayOut {
  Strided x;
  __device__ __forceinline__ void operator()(int32_t, int32_t o, float v) const { x(o) = v; }
};

// NaN-propagating max/min with the semantics of jnp.maximum / jnp.minimum.
// Select form: both operands are always used, so the compiler cannot sink the
// computation of one of them (a load, say) into a branch.
__device__ __forceinline__ float nan_max(float a, float b) {
  float r = fmaxf(a, b);
  r = (b != b) ? b : r;
  return (a != a) ? a : r;
}
__device__ __forceinline__ float nan_min(float a, float b) {
  float r = fminf(a, b);
  r = (b != b) ? b : r;
  return (a != a) ? a : r;
}

}  // namespace wmpc
