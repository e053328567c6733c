// Device-side building blocks of the batched fixed-pattern LDLᵀ (gfx950).
//
// One lane owns one instance.  Every per-instance array is batch-major
// ([entry][ld]), so a wavefront touching entry i of 64 consecutive instances
// issues one contiguous 256-byte access.  All table reads are wave-uniform and
// compile to scalar (SMEM) loads.
//
// The row program follows the up-looking QDLDL factorization of
// warpmpc/jax_qdldl/core.py:867-933 (== warp_backend.py:104-173); the dense
// scratch vector y of the reference is replaced by the R <= max_reach slots of
// the reach of the current row, kept in LDS.
#pragma once

#include "wmpc_common.h"

namespace wmpc {

// Values straight from a [nnz_a][ld] array (CompiledQDLDL.factor input).
struct DirectSource {
  const float* __restrict__ a;
  int64_t ld;
  __device__ __forceinline__ float operator()(int32_t slot, int64_t b) const {
    return a[(int64_t)slot * ld + b];
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
    const int64_t i = pm_ld ? (int64_t)idx[slot] * pm_ld + b : (int64_t)idx[slot];
    if (k == 0) return p[i] * scale[slot] + off[slot];
    if (k == 1) return a[i] * scale[slot];
    return off[slot];
  }
};

// Factor one instance: values of source instance sb, outputs at instance b.
// y: this lane's LDS slot column (stride ys floats).  d may be null.
// FUSE_FWD additionally runs the row-oriented forward substitution
// w[k] = rhs_perm[k] - sum_c L[k,c] w[c] with the freshly computed L[k, :].
template <class Src, bool FUSE_FWD, class Rhs>
__device__ __forceinline__ void factor_lane(const QdldlTables& t, const Src& src, int64_t sb,
                                            int64_t b, int64_t ld, float* __restrict__ lx,
                                            float* __restrict__ d, float* __restrict__ dinv,
                                            float* y, int ys, const Rhs& rhs,
                                            float* __restrict__ w) {
  for (int32_t k = 0; k < t.n; ++k) {
    const int32_t r0 = t.proc_ptr[k];
    const int32_t r1 = t.proc_ptr[k + 1];
    for (int32_t s = 0; s < r1 - r0; ++s) y[s * ys] = 0.0f;
    for (int32_t j = t.init_ptr[k]; j < t.init_ptr[k + 1]; ++j)
      y[t.init_slot[j] * ys] += src(t.init_src[j], sb);
    float dk = src(t.diag_src[k], sb);
    float fw = 0.0f;
    if constexpr (FUSE_FWD) fw = rhs(t.p[k], b);
    for (int32_t r = r0; r < r1; ++r) {
      const int32_t c = t.proc_col[r];
      const int32_t u0 = t.upd_ptr[r];
      const int32_t u1 = t.upd_ptr[r + 1];
      const int32_t base = t.lp[c] - u0;
      const float val = y[(r - r0) * ys];
      for (int32_t u = u0; u < u1; ++u) {
        const float l = lx[(int64_t)(base + u) * ld + b];
        y[t.upd_slot[u] * ys] -= l * val;
      }
      const float lval = val * dinv[(int64_t)c * ld + b];
      lx[(int64_t)t.proc_lidx[r] * ld + b] = lval;
      dk -= val * lval;
      if constexpr (FUSE_FWD) fw -= lval * w[(int64_t)c * ld + b];
    }
    if (d) d[(int64_t)k * ld + b] = dk;
    dinv[(int64_t)k * ld + b] = 1.0f / dk;
    if constexpr (FUSE_FWD) w[(int64_t)k * ld + b] = fw;
  }
}

// Forward substitution by rows (gather form): w[k] = rhs_perm[k] - sum L[k,c] w[c].
// The entries of row k of L are exactly the proc entries of row k.
// SHARED: one factor for the whole batch, stored as contiguous [nnz_l] / [n].
template <bool SHARED, class Rhs>
__device__ __forceinline__ void forward_rows(const QdldlTables& t, const float* __restrict__ lx,
                                             int64_t f_ld, int64_t b, int64_t ld, const Rhs& rhs,
                                             float* __restrict__ w) {
  const int64_t fb = SHARED ? 0 : b;
  const int64_t fs = SHARED ? 1 : f_ld;
  for (int32_t k = 0; k < t.n; ++k) {
    float acc = rhs(t.p[k], b);
    for (int32_t r = t.proc_ptr[k]; r < t.proc_ptr[k + 1]; ++r) {
      const float l = lx[(int64_t)t.proc_lidx[r] * fs + fb];
      acc -= l * w[(int64_t)t.proc_col[r] * ld + b];
    }
    w[(int64_t)k * ld + b] = acc;
  }
}

// D⁻¹ scaling fused with the backward column sweep (core.py:1003-1028):
// w[c] = w[c] * dinv[c] - sum_q L[q] w[li[q]], c = n-1 .. 0.
template <bool SHARED>
__device__ __forceinline__ void backward_cols(const QdldlTables& t, const float* __restrict__ lx,
                                              const float* __restrict__ dinv, int64_t f_ld,
                                              int64_t b, int64_t ld, float* __restrict__ w) {
  const int64_t fb = SHARED ? 0 : b;
  const int64_t fs = SHARED ? 1 : f_ld;
  for (int32_t c = t.n - 1; c >= 0; --c) {
    float corr = 0.0f;
    for (int32_t q = t.lp[c]; q < t.lp[c + 1]; ++q)
      corr += lx[(int64_t)q * fs + fb] * w[(int64_t)t.li[q] * ld + b];
    w[(int64_t)c * ld + b] = w[(int64_t)c * ld + b] * dinv[(int64_t)c * fs + fb] - corr;
  }
}

struct ArrayRhs {
  const float* __restrict__ rhs;
  int64_t ld;
  __device__ __forceinline__ float operator()(int32_t i, int64_t b) const {
    return rhs[(int64_t)i * ld + b];
  }
};

}  // namespace wmpc
