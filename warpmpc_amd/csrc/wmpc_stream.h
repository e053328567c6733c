// Chunked, prefetched column-oriented triangular sweeps (gfx950).
//
// With one lane per instance a batch of 65,536 fills the chip with exactly one
// wavefront per SIMD, so latency cannot be hidden by other wavefronts: each
// wavefront must keep many loads in flight itself.  A sweep therefore walks a
// precomputed record stream (column markers and L entries, in chunks of 64
// records) and every operand it reads is a sequential stream consumed chunk by
// chunk: the L values of the instance (or of the shared factor), the staged
// right-hand side / stash values and D⁻¹.  Each chunk's operands are loaded
// into registers two chunks before use, then committed to LDS (per-lane
// columns in [slot][64] layout, conflict-free; wave-uniform data in a shared
// area read by broadcast).  Every chunk issues the same loads unconditionally,
// so the compiler's memory wait counts stay exact and the prefetch is real.
#pragma once

#include "wmpc_qdldl_device.h"

namespace wmpc {

// Plain element types of the wave streams (HIP's int4/int2 are union-based
// vector types that the compiler cannot split into registers here).
struct I4 {
  int32_t x, y, z, w;
};
struct I2 {
  int32_t x, y;
};

// Global-memory pointers.  Sweep operands are addressed through these so the
// loads are global_load_* (vector memory only): a generic pointer would give
// flat loads, which also count against the LDS wait counter and force full
// waits before every LDS access.
#define WMPC_GLOBAL __attribute__((address_space(1)))
template <class T>
using gptr = WMPC_GLOBAL T*;

// A wave-uniform global pointer the compiler must treat as unknown at each
// call: keeps the address arithmetic of a sweep's set-up inside the sweep
// instead of being hoisted out of the caller's iteration loop (where the
// hoisted addresses would stay live across every iteration and spill).
template <class T>
__device__ __forceinline__ gptr<T> opaque(const T* p) {
  gptr<T> g = (gptr<T>)(p);
  asm volatile("" : "+s"(g));
  return g;
}

// Device view of the chunked record streams (built in wmpc_program.h).
struct RecStream {
  const I2* rec = nullptr;  // [nchunks * kChunkRecs]
  int32_t nchunks = 0;
  int32_t l1 = 0, v1 = 0, d1 = 0;  // stream offsets of chunk 1
};
struct SweepProg {
  int32_t n = 0, nnz_l = 0, nw = 0;
  int32_t lcols = 64;  // LDS columns of the row window (QdldlProg::lcols)
  RecStream fwd, bwd;
};

constexpr int kChunkRecs = 64;       // records per chunk (2 header + 62)
constexpr int32_t kMarkBit = 1 << 20;
constexpr int32_t kSlotMask = 0xFFFF;
constexpr int32_t kNoSlot = 0xFFFF;
constexpr int32_t kFirstTouchBit = 1 << 16;
constexpr int kVPer = 16;            // per-column values per chunk
constexpr int kLPer = 64;            // L values per chunk

__host__ __device__ inline SweepProg sweep_prog(const QdldlProg& P) {
  SweepProg S;
  S.n = P.n;
  S.nnz_l = P.nnz_l;
  S.nw = P.nbw;
  S.lcols = P.lcols;
  S.fwd = RecStream{reinterpret_cast<const I2*>(P.frec), P.fchunks, P.f1[0], P.f1[1], P.f1[2]};
  S.bwd = RecStream{reinterpret_cast<const I2*>(P.brec), P.bchunks, P.b1[0], P.b1[1], P.b1[2]};
  return S;
}

// LDS floats of the sweeps in a workgroup: the W window ([slot][lcols]), then
// the current chunk's L values of a per-instance factor and its per-column
// values, one column per lane ([slot][64], constant offsets); wave-uniform
// chunk data stays in registers.
__host__ __device__ inline int64_t sweep_lds_floats(const SweepProg& S, bool shared_l) {
  return (int64_t)S.nw * S.lcols + (int64_t)(shared_l ? kVPer : kLPer + 2 * kVPer) * kWave;
}

// One chunk of a sweep in flight: loaded into registers two chunks ahead of
// use.  Every chunk issues the same loads, so the compiler's wait counts
// before the chunk is taken over stay exact.
//   rec, lu, du: wave-uniform data, lane j holding element j (record j, the
//                shared L value and D⁻¹ value j of the chunk); read back with
//                readlane into scalar registers;
//   l, v       : per-lane L values (per-instance factor) and per-column value
//                streams, committed to the lane's LDS column.
template <bool SHL, int NV>
struct Chunk {
  I2 rec;
  float lu, du;
  float l[SHL ? 1 : kLPer];
  float v[NV][kVPer];
  int32_t lbase, vbase, dbase;
};

__device__ __forceinline__ int32_t rl(int32_t v, int32_t k) {
  return __builtin_amdgcn_readlane(v, k);
}
__device__ __forceinline__ float rlf(float v, int32_t k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

template <bool SHL, int NV>
struct Sweep {
  float* W;      // lane column: row window, slot s at W[s * lc]
  float* lring;  // lane column: L values of the current chunk (per-instance factor)
  float* vring;  // lane column: NV value streams of the current chunk
  int lane;
  int32_t lc;
  // the current chunk's wave-uniform data
  int32_t crx, cry;
  float clu, cdu;

  __device__ __forceinline__ void setup(float* lds_base, int32_t nw, int32_t lc_, int lane_) {
    lane = lane_;
    lc = lc_;
    W = lds_base + (lane_ & (lc_ - 1));
    lring = lds_base + (int64_t)nw * lc_ + lane_;
    vring = lring + (SHL ? 0 : kLPer * kWave);
  }

  template <class LSrc, class V0Src, class V1Src, class DSrc>
  __device__ __forceinline__ void issue(Chunk<SHL, NV>& c, gptr<const I2> rec, int32_t t, int32_t lb,
                                        int32_t vb, int32_t db, const LSrc& ls, const V0Src& v0,
                                        const V1Src& v1, const DSrc& ds) {
    const int64_t ri = (int64_t)t * kChunkRecs + lane;
    c.rec.x = rec[ri].x;  // member-wise: the record lives in global memory
    c.rec.y = rec[ri].y;
    c.lbase = lb;
    c.vbase = vb;
    c.dbase = db;
    if constexpr (SHL) {
      c.lu = ls(lb + lane);
      c.du = ds(db + lane);
    } else {
#pragma unroll
      for (int j = 0; j < kLPer; ++j) c.l[j] = ls(lb + j);
    }
#pragma unroll
    for (int j = 0; j < kVPer; ++j) c.v[0][j] = v0(vb + j);
    if constexpr (NV > 1) {
#pragma unroll
      for (int j = 0; j < kVPer; ++j) c.v[1][j] = v1(db + j);
    }
  }
  __device__ __forceinline__ void take(const Chunk<SHL, NV>& c) {
    crx = c.rec.x;
    cry = c.rec.y;
    if constexpr (SHL) {
      clu = c.lu;
      cdu = c.du;
    } else {
#pragma unroll
      for (int j = 0; j < kLPer; ++j) lring[j * kWave] = c.l[j];
    }
#pragma unroll
    for (int j = 0; j < kVPer; ++j) vring[j * kWave] = c.v[0][j];
    if constexpr (NV > 1) {
#pragma unroll
      for (int j = 0; j < kVPer; ++j) vring[(kVPer + j) * kWave] = c.v[1][j];
    }
  }
  __device__ __forceinline__ int32_t rec_w(int32_t k) const { return rl(crx, k); }
  __device__ __forceinline__ int32_t rec_a(int32_t k) const { return rl(cry, k); }
  __device__ __forceinline__ float L(int32_t j) const {
    if constexpr (SHL) return rlf(clu, j);
    return lring[j * kWave];
  }
  __device__ __forceinline__ float V(int k, int32_t j) const { return vring[(k * kVPer + j) * kWave]; }
  __device__ __forceinline__ float D(int32_t j) const {
    if constexpr (SHL) return rlf(cdu, j);
    return vring[(kVPer + j) * kWave];
  }
  // number of consecutive entry records starting at k (k is an entry), at most G
  template <int G>
  __device__ __forceinline__ int32_t run(int32_t k, int32_t kend) const {
    int32_t g = 1;
    while (g < G && k + g < kend && !(rec_w(k + g) & kMarkBit)) ++g;
    return g;
  }
};

// G entries of one forward column: distinct rows, all window reads first;
// first updates start from the staged right-hand side.  Returns the number of
// right-hand-side values consumed.
template <int G, bool SHL, int NV>
__device__ __forceinline__ int32_t fwd_group(Sweep<SHL, NV>& sw, int32_t k, int32_t j, int32_t v,
                                             float val) {
  int32_t sl[G], f[G];
  float l[G], w[G];
  int32_t fa = 0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int32_t rw = sw.rec_w(k + g);
    sl[g] = rw & kSlotMask;
    f[g] = (rw >> 16) & 1;
    fa |= f[g];
    l[g] = sw.L(j + g);
  }
#pragma unroll
  for (int g = 0; g < G; ++g) w[g] = sw.W[sl[g] * sw.lc];
  int32_t used = 0;
  if (fa) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (f[g]) {
        w[g] = sw.V(0, v + used);
        ++used;
      }
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) sw.W[sl[g] * sw.lc] = w[g] - l[g] * val;
  return used;
}

template <int G, bool SHL, int NV>
__device__ __forceinline__ float bwd_group(const Sweep<SHL, NV>& sw, int32_t k, int32_t j) {
  int32_t sl[G];
  float l[G], w[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    sl[g] = sw.rec_w(k + g) & kSlotMask;
    l[g] = sw.L(j + g);
  }
#pragma unroll
  for (int g = 0; g < G; ++g) w[g] = sw.W[sl[g] * sw.lc];
  float acc = 0.f;
#pragma unroll
  for (int g = 0; g < G; ++g) acc += l[g] * w[g];
  return acc;
}

// Clamped sources.
struct LaneSrc {  // ptr[i * ld + b], i clamped to [0, count)
  gptr<const float> ptr;
  int64_t ld, b;
  int32_t count;
  __device__ __forceinline__ float operator()(int32_t i) const {
    i = i < count ? i : count - 1;
    return (ptr + (int64_t)i * ld)[(uint32_t)b];
  }
};
struct LaneSrcRev {  // ptr[(count - 1 - i) * ld + b]
  gptr<const float> ptr;
  int64_t ld, b;
  int32_t count;
  __device__ __forceinline__ float operator()(int32_t i) const {
    i = i < count ? i : count - 1;
    return (ptr + (int64_t)(count - 1 - i) * ld)[(uint32_t)b];
  }
};
struct NoSrc {
  __device__ __forceinline__ float operator()(int32_t) const { return 0.f; }
};

// Forward substitution, column (scatter) form, over the forward record
// stream: a marker finalises row c (stash(c) = w[c] * D⁻¹[c]); entries apply
// w[li[q]] -= L[q] w[c], a row's first update starting from its right-hand
// side (rhs in first-touch order, rhs[i * ld + b]).
// SHL: one factor shared by the batch (lx [nnz_l], dinv [n]).
template <bool SHL>
__device__ __forceinline__ void fwd_cols(const SweepProg& S, const float* __restrict__ lx,
                                         const float* __restrict__ dinv, int64_t f_ld, int64_t fb,
                                         const float* __restrict__ rhs, float* __restrict__ stash,
                                         int64_t ld, int64_t b, float* lds_base, int lane) {
  constexpr int NV = SHL ? 1 : 2;
  Sweep<SHL, NV> sw;
  sw.setup(lds_base, S.nw, S.lcols, lane);
  const RecStream R = S.fwd;
  const gptr<const I2> rec = opaque(R.rec);
  const gptr<float> st = (gptr<float>)opaque(stash);
  const int32_t nl = S.nnz_l > 0 ? S.nnz_l : 1;
  const LaneSrc ls{opaque(lx), SHL ? 1 : f_ld, SHL ? 0 : fb, nl};
  const LaneSrc ds{opaque(dinv), SHL ? 1 : f_ld, SHL ? 0 : fb, S.n};
  const LaneSrc rs{opaque(rhs), ld, b, S.n};
  Chunk<SHL, NV> ca, cb;
  sw.issue(ca, rec, 0, 0, 0, 0, ls, rs, ds, ds);
  sw.issue(cb, rec, 1, R.l1, R.v1, R.d1, ls, rs, ds, ds);
  float val = 0.f;
  int32_t li = 0, vi = 0, di = 0, col = 0;
  // Each record step computes its counter increments and applies them once at
  // the end (branch arms that end in the same increments would otherwise be
  // merged through pointers to the counters, pinning them in scratch memory).
  auto process = [&](int32_t lb, int32_t vb, int32_t db) __attribute__((always_inline)) {
    const int32_t kend = 2 + sw.rec_a(1);
    int32_t k = 2;
    while (k < kend) {
      const int32_t rw = sw.rec_w(k);
      int32_t dk = 1, dl = 0, dv = 0, dd = 0;
      if (rw & kMarkBit) {
        const int32_t slot = rw & kSlotMask;
        if (slot != kNoSlot) {
          val = sw.W[slot * sw.lc];
        } else {
          val = sw.V(0, vi - vb);
          dv = 1;
        }
        const float dval = sw.D(di - db);
        dd = 1;
        (st + (int64_t)col * ld)[(uint32_t)b] = val * dval;
        ++col;
      } else {
        const int32_t g = sw.template run<8>(k, kend);
        if (g == 8) {
          dv = fwd_group<8>(sw, k, li - lb, vi - vb, val);
          dk = dl = 8;
        } else if (g >= 4) {
          dv = fwd_group<4>(sw, k, li - lb, vi - vb, val);
          dk = dl = 4;
        } else {
          dv = fwd_group<1>(sw, k, li - lb, vi - vb, val);
          dk = dl = 1;
        }
      }
      k += dk;
      li += dl;
      vi += dv;
      di += dd;
    }
  };
  // rotated pipeline: every take-over of a chunk sits inside the loop body,
  // where the wait for its loads is counted exactly (at a loop header the
  // compiler would wait for all outstanding loads)
  auto next = [&](Chunk<SHL, NV>& c, int32_t t) __attribute__((always_inline)) {
    sw.take(c);
    const int32_t lb = c.lbase, vb = c.vbase, db = c.dbase;
    const int32_t l2 = sw.rec_w(0), v2 = sw.rec_a(0), d2 = sw.rec_w(1);
    sw.issue(c, rec, t < R.nchunks ? t : R.nchunks - 1, l2, v2, d2, ls, rs, ds, ds);
    return I4{lb, vb, db, 0};
  };
  I4 base = next(ca, 2);
  for (int32_t t = 0; t < R.nchunks; t += 2) {
    process(base.x, base.y, base.z);
    base = next(cb, t + 3);
    process(base.x, base.y, base.z);
    base = next(ca, t + 4);
  }
}

// Backward substitution, column (gather) form, over the backward record
// stream: entries accumulate sum L[q] w[li[q]]; the marker of column c sets
// w[c] = stash(c) - sum (stash already D⁻¹-scaled) and hands it to
// out(c, p[c], w[c]).
template <bool SHL, class Out>
__device__ __forceinline__ void bwd_cols(const SweepProg& S, const float* __restrict__ lx,
                                         int64_t f_ld, int64_t fb, const float* __restrict__ stash,
                                         int64_t ld, int64_t b, float* lds_base, int lane,
                                         const Out& out) {
  Sweep<SHL, 1> sw;
  sw.setup(lds_base, S.nw, S.lcols, lane);
  const RecStream R = S.bwd;
  const gptr<const I2> rec = opaque(R.rec);
  const int32_t nl = S.nnz_l > 0 ? S.nnz_l : 1;
  const LaneSrcRev ls{opaque(lx), SHL ? 1 : f_ld, SHL ? 0 : fb, nl};
  const LaneSrcRev ss{opaque(stash), ld, b, S.n};
  const NoSrc none{};
  Chunk<SHL, 1> ca, cb;
  sw.issue(ca, rec, 0, 0, 0, 0, ls, ss, none, none);
  sw.issue(cb, rec, 1, R.l1, R.v1, R.d1, ls, ss, none, none);
  float corr = 0.f;
  int32_t li = 0, vi = 0, col = S.n - 1;
  auto process = [&](int32_t lb, int32_t vb) __attribute__((always_inline)) {
    const int32_t kend = 2 + sw.rec_a(1);
    int32_t k = 2;
    while (k < kend) {
      const int32_t rw = sw.rec_w(k);
      int32_t dk = 1, dl = 0, dv = 0;
      if (rw & kMarkBit) {
        const float w = sw.V(0, vi - vb) - corr;
        dv = 1;
        corr = 0.f;
        const int32_t slot = rw & kSlotMask;
        if (slot != kNoSlot) sw.W[slot * sw.lc] = w;
        out(col, sw.rec_a(k), w);
        --col;
      } else {
        const int32_t g = sw.template run<8>(k, kend);
        if (g == 8) {
          corr += bwd_group<8>(sw, k, li - lb);
          dk = dl = 8;
        } else if (g >= 4) {
          corr += bwd_group<4>(sw, k, li - lb);
          dk = dl = 4;
        } else {
          corr += bwd_group<1>(sw, k, li - lb);
          dk = dl = 1;
        }
      }
      k += dk;
      li += dl;
      vi += dv;
    }
  };
  auto next = [&](Chunk<SHL, 1>& c, int32_t t) __attribute__((always_inline)) {
    sw.take(c);
    const int32_t lb = c.lbase, vb = c.vbase;
    const int32_t l2 = sw.rec_w(0), v2 = sw.rec_a(0);
    sw.issue(c, rec, t < R.nchunks ? t : R.nchunks - 1, l2, v2, 0, ls, ss, none, none);
    return I2{lb, vb};
  };
  I2 base = next(ca, 2);
  for (int32_t t = 0; t < R.nchunks; t += 2) {
    process(base.x, base.y);
    base = next(cb, t + 3);
    process(base.x, base.y);
    base = next(ca, t + 4);
  }
}

}  // namespace wmpc
