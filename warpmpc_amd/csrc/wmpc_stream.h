// Prefetched streams and the column-oriented triangular sweeps (gfx950).
//
// With one lane per instance a batch of 65,536 fills the chip with exactly one
// wavefront per SIMD, so latency cannot be hidden by other wavefronts: each
// wavefront must keep many loads in flight itself.  Every operand the sweeps
// read is therefore consumed as a sequential stream and prefetched:
//
//   * LaneStream — a per-lane float stream (L values of one instance, the
//     right-hand side, the forward result).  Two chunks of CH values are in
//     flight in VGPRs while two more chunks sit in an LDS ring of the lane's
//     own column ([slot][64] layout, conflict-free);
//   * WaveStream — a wave-uniform stream (program tables, a shared factor).
//     Lane j loads element j of a chunk; the chunk is staged in a wave-shared
//     LDS ring and read back with a uniform address (an LDS broadcast).
//
// A stream element i is readable once the consumer has advanced to it; the
// consumer reads indices in increasing order, and crossing into a new chunk
// commits the next staged chunk and issues the loads of the chunk after it.
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

template <int CH>
struct LaneStream {
  static_assert((CH & (CH - 1)) == 0, "chunk must be a power of two");
  static constexpr int kSlots = 2 * CH;  // LDS slots of the ring
  float r0[CH], r1[CH];
  float* ring;
  int64_t cur;
  int64_t count;

  template <class Src>
  __device__ __forceinline__ void fill(float (&r)[CH], int64_t chunk, const Src& src) {
    const int64_t base = chunk * CH;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int64_t i = base + j;
      r[j] = src(i < count ? i : count - 1);
    }
  }
  __device__ __forceinline__ void commit(const float (&r)[CH], int half) {
#pragma unroll
    for (int j = 0; j < CH; ++j) ring[(half * CH + j) * kWave] = r[j];
  }
  template <class Src>
  __device__ __forceinline__ void init(float* ring_, int64_t count_, const Src& src) {
    ring = ring_;
    count = count_ > 0 ? count_ : 1;
    cur = 0;
    fill(r0, 0, src);
    fill(r1, 1, src);
    commit(r0, 0);
    fill(r0, 2, src);
    commit(r1, 1);
    fill(r1, 3, src);
  }
  // Element i readable with at(i) for every i in [current, i]; crosses at most
  // one chunk boundary, so a group read spans fewer than CH elements.
  template <class Src>
  __device__ __forceinline__ void ensure(int64_t i, const Src& src) {
    (void)get(i, src);
  }
  __device__ __forceinline__ float at(int64_t i) const {
    return ring[(((i / CH) & 1) * CH + (i & (CH - 1))) * kWave];
  }
  template <class Src>
  __device__ __forceinline__ float get(int64_t i, const Src& src) {
    if (i >= (cur + 2) * CH) {
      // chunk cur is consumed: stage chunk cur+2 in its half, load chunk cur+4.
      // The two arms differ only in which register set they touch; the
      // markers at both ends keep the compiler from hoisting or sinking their
      // common code into one path that selects the set through a pointer
      // (which would demote both sets to scratch memory).
      if (cur & 1) {
        asm volatile("; stream set 1 begin");
        commit(r1, 1);
        fill(r1, cur + 4, src);
        asm volatile("; stream set 1 end");
      } else {
        asm volatile("; stream set 0 begin");
        commit(r0, 0);
        fill(r0, cur + 4, src);
        asm volatile("; stream set 0 end");
      }
      ++cur;
    }
    return ring[(((i / CH) & 1) * CH + (i & (CH - 1))) * kWave];
  }
};

// Wave-uniform stream of T (4 or 8 or 16 bytes) with 64-element chunks.
template <class T>
struct WaveStream {
  static constexpr int CH = kWave;
  static constexpr int kElems = 2 * CH;  // ring elements (wave-shared)
  T r0, r1;
  T* ring;
  int64_t cur;
  int64_t count;
  int lane;

  template <class Src>
  __device__ __forceinline__ T load(int64_t chunk, const Src& src) const {
    const int64_t i = chunk * CH + lane;
    return src(i < count ? i : count - 1);
  }
  __device__ __forceinline__ void commit(const T& r, int half) {
    ring[half * CH + lane] = r;
    __builtin_amdgcn_wave_barrier();
  }
  template <class Src>
  __device__ __forceinline__ void init(T* ring_, int64_t count_, int lane_, const Src& src) {
    ring = ring_;
    count = count_ > 0 ? count_ : 1;
    lane = lane_;
    cur = 0;
    r0 = load(0, src);
    r1 = load(1, src);
    commit(r0, 0);
    r0 = load(2, src);
    commit(r1, 1);
    r1 = load(3, src);
  }
  template <class Src>
  __device__ __forceinline__ void ensure(int64_t i, const Src& src) {
    (void)get(i, src);
  }
  __device__ __forceinline__ T at(int64_t i) const {
    return ring[((i / CH) & 1) * CH + (i & (CH - 1))];
  }
  template <class Src>
  __device__ __forceinline__ T get(int64_t i, const Src& src) {
    if (i >= (cur + 2) * CH) {
      if (cur & 1) {
        asm volatile("; wave stream set 1 begin");
        commit(r1, 1);
        r1 = load(cur + 4, src);
        asm volatile("; wave stream set 1 end");
      } else {
        asm volatile("; wave stream set 0 begin");
        commit(r0, 0);
        r0 = load(cur + 4, src);
        asm volatile("; wave stream set 0 end");
      }
      ++cur;
    }
    return ring[((i / CH) & 1) * CH + (i & (CH - 1))];
  }
};

// A wave-uniform pointer the compiler must treat as unknown at each call:
// keeps the address arithmetic of a sweep's stream set-up inside the sweep
// instead of being hoisted out of the caller's iteration loop (where the
// hoisted addresses would stay live across every iteration and spill).
template <class T>
__device__ __forceinline__ T* opaque(T* p) {
  asm volatile("" : "+s"(p));
  return p;
}

// Sources of the streams.
struct LaneArray {  // ptr[i * ld + b]: uniform row base + 32-bit lane offset
  const float* __restrict__ ptr;
  int64_t ld, b;
  __device__ __forceinline__ float operator()(int64_t i) const {
    return (ptr + i * ld)[(uint32_t)b];
  }
};
struct LaneArrayRev {  // ptr[(n - 1 - i) * ld + b]
  const float* __restrict__ ptr;
  int64_t ld, b, n;
  __device__ __forceinline__ float operator()(int64_t i) const {
    return (ptr + (n - 1 - i) * ld)[(uint32_t)b];
  }
};

// Device tables of the column sweeps (see wmpc_program.h).
struct SweepProg {
  int32_t n = 0, nnz_l = 0, nw = 0;
  const I4* fcol = nullptr;      // [n]   {W slot of row c or -1, lp[c+1], 0, 0}
  const I4* bcol = nullptr;      // [n]   reverse: {W slot of row c or -1, lp[c], p[c], 0}
  const int32_t* fent = nullptr; // [nnz_l] W slot of li[q] | kFirstTouch
};

constexpr int32_t kFirstTouchBit = 1 << 16;
constexpr int kLChunk = 32;   // L values per lane chunk
constexpr int kVChunk = 8;    // per-column values per lane chunk

__host__ __device__ inline SweepProg sweep_prog(const QdldlProg& P) {
  SweepProg S;
  S.n = P.n;
  S.nnz_l = P.nnz_l;
  S.nw = P.nbw;
  S.fcol = reinterpret_cast<const I4*>(P.fcol);
  S.bcol = reinterpret_cast<const I4*>(P.bcol2);
  S.fent = P.fent;
  return S;
}

// LDS floats per lane of the sweeps (W window, L ring unless the factor is
// shared, right-hand-side / D⁻¹ / stash rings), and the bytes of the
// wave-shared rings.
__host__ __device__ inline int32_t sweep_lane_slots(const SweepProg& S, bool shared_l) {
  return S.nw + (shared_l ? 2 * kVChunk : 2 * kLChunk + 4 * kVChunk);
}
__host__ __device__ inline int32_t sweep_wave_bytes() {
  return 2 * kWave * (int32_t)(sizeof(I4) + sizeof(I2));
}

// Forward substitution, column (scatter) form, c = 0 .. n-1:
//   w[c] is final at column c; stash(c) = w[c] * dinv(c);
//   w[li[q]] -= L[q] w[c] for the entries of column c, the first update of a
//   row starting from its right-hand side.
// rhs: right-hand side in first-touch order, R(i) at rhs[i * ld + b].
// SHL: one factor shared by the batch (L, dinv wave-uniform).
template <bool SHL>
__device__ __forceinline__ void fwd_cols(const SweepProg& S, const float* __restrict__ lx,
                                         const float* __restrict__ dinv, int64_t f_ld, int64_t fb,
                                         const float* __restrict__ rhs, float* __restrict__ stash,
                                         int64_t ld, int64_t b, float* lane_lds, char* wave_lds,
                                         int lane) {
  float* W = lane_lds;
  float* lring = W + S.nw * kWave;
  float* vring = lring + (SHL ? 0 : 2 * kLChunk * kWave);
  I4* cring = reinterpret_cast<I4*>(wave_lds);
  I2* ering = reinterpret_cast<I2*>(cring + 2 * kWave);
  const I4* fcol = opaque(S.fcol);
  const int32_t* fent = opaque(S.fent);
  lx = opaque(lx);
  dinv = opaque(dinv);
  rhs = opaque(rhs);
  stash = opaque(stash);
  // column info (+ the shared dinv), entry info (+ the shared L value)
  auto csrc = [&](int64_t c) {
    I4 v = fcol[c];
    if constexpr (SHL) v.z = __float_as_int(dinv[c]);
    return v;
  };
  auto esrc = [&](int64_t q) {
    I2 v;
    v.x = fent[q];
    v.y = SHL ? __float_as_int(lx[q]) : 0;
    return v;
  };
  const LaneArray lsrc{lx, f_ld, fb};
  const LaneArray dsrc{dinv, f_ld, fb};
  const LaneArray rsrc{rhs, ld, b};
  WaveStream<I4> C;
  WaveStream<I2> E;
  LaneStream<kLChunk> Ls;
  LaneStream<kVChunk> Rs;
  LaneStream<kVChunk> Ds;
  C.init(cring, S.n, lane, csrc);
  if (S.nnz_l > 0) E.init(ering, S.nnz_l, lane, esrc);
  if constexpr (!SHL) {
    if (S.nnz_l > 0) Ls.init(lring, S.nnz_l, lsrc);
    Ds.init(vring + 2 * kVChunk * kWave, S.n, dsrc);
  }
  Rs.init(vring, S.n, rsrc);
  int64_t q = 0, ir = 0;
  for (int32_t c = 0; c < S.n; ++c) {
    const I4 ci = C.get(c, csrc);
    const float val = ci.x >= 0 ? W[ci.x * kWave] : Rs.get(ir++, rsrc);
    const float di = SHL ? __int_as_float(ci.z) : Ds.get(c, dsrc);
    (stash + (int64_t)c * ld)[(uint32_t)b] = val * di;
    // entries of one column update distinct rows: groups of four issue all
    // their LDS reads before any write (first updates start from the staged
    // right-hand side instead of the window)
    for (; q + 4 <= ci.y; q += 4) {
      E.ensure(q + 3, esrc);
      if constexpr (!SHL) Ls.ensure(q + 3, lsrc);
      Rs.ensure(ir + 3, rsrc);
      const I2 e0 = E.at(q), e1 = E.at(q + 1), e2 = E.at(q + 2), e3 = E.at(q + 3);
      const float l0 = SHL ? __int_as_float(e0.y) : Ls.at(q);
      const float l1 = SHL ? __int_as_float(e1.y) : Ls.at(q + 1);
      const float l2 = SHL ? __int_as_float(e2.y) : Ls.at(q + 2);
      const float l3 = SHL ? __int_as_float(e3.y) : Ls.at(q + 3);
      const int32_t s0 = e0.x & (kFirstTouchBit - 1), s1 = e1.x & (kFirstTouchBit - 1);
      const int32_t s2 = e2.x & (kFirstTouchBit - 1), s3 = e3.x & (kFirstTouchBit - 1);
      const int32_t f0 = (e0.x >> 16) & 1, f1 = (e1.x >> 16) & 1;
      const int32_t f2 = (e2.x >> 16) & 1, f3 = (e3.x >> 16) & 1;
      const float w0 = W[s0 * kWave], w1 = W[s1 * kWave], w2 = W[s2 * kWave], w3 = W[s3 * kWave];
      const float r0 = Rs.at(ir), r1 = Rs.at(ir + f0), r2 = Rs.at(ir + f0 + f1);
      const float r3 = Rs.at(ir + f0 + f1 + f2);
      W[s0 * kWave] = (f0 ? r0 : w0) - l0 * val;
      W[s1 * kWave] = (f1 ? r1 : w1) - l1 * val;
      W[s2 * kWave] = (f2 ? r2 : w2) - l2 * val;
      W[s3 * kWave] = (f3 ? r3 : w3) - l3 * val;
      ir += f0 + f1 + f2 + f3;
    }
    for (; q < ci.y; ++q) {
      const I2 e = E.get(q, esrc);
      const float l = SHL ? __int_as_float(e.y) : Ls.get(q, lsrc);
      const int32_t slot = e.x & (kFirstTouchBit - 1);
      if (e.x & kFirstTouchBit) {
        W[slot * kWave] = Rs.get(ir++, rsrc) - l * val;
      } else {
        W[slot * kWave] -= l * val;
      }
    }
  }
}

// Backward substitution, column (gather) form, c = n-1 .. 0:
//   w[c] = stash(c) - sum_q L[q] w[li[q]]   (stash already D⁻¹-scaled);
// out(c, p[c], w[c]) receives the solution entry of original index p[c].
template <bool SHL, class Out>
__device__ __forceinline__ void bwd_cols(const SweepProg& S, const float* __restrict__ lx,
                                         int64_t f_ld, int64_t fb, const float* __restrict__ stash,
                                         int64_t ld, int64_t b, float* lane_lds, char* wave_lds,
                                         int lane, const Out& out) {
  float* W = lane_lds;
  float* lring = W + S.nw * kWave;
  float* vring = lring + (SHL ? 0 : 2 * kLChunk * kWave);
  I4* cring = reinterpret_cast<I4*>(wave_lds);
  I2* ering = reinterpret_cast<I2*>(cring + 2 * kWave);
  const I4* bcol = opaque(S.bcol);
  const int32_t* fent = opaque(S.fent);
  lx = opaque(lx);
  stash = opaque(stash);
  const int64_t nl = S.nnz_l;
  auto csrc = [&](int64_t i) { return bcol[i]; };
  // entries in descending q
  auto esrc = [&](int64_t i) {
    I2 v;
    v.x = fent[nl - 1 - i];
    v.y = SHL ? __float_as_int(lx[nl - 1 - i]) : 0;
    return v;
  };
  const LaneArrayRev lsrc{lx, f_ld, fb, nl};
  const LaneArrayRev ssrc{stash, ld, b, S.n};
  WaveStream<I4> C;
  WaveStream<I2> E;
  LaneStream<kLChunk> Ls;
  LaneStream<kVChunk> Ss;
  C.init(cring, S.n, lane, csrc);
  if (nl > 0) E.init(ering, nl, lane, esrc);
  if constexpr (!SHL) {
    if (nl > 0) Ls.init(lring, nl, lsrc);
  }
  Ss.init(vring, S.n, ssrc);
  int64_t i = 0;  // entries consumed (q = nl - 1 - i)
  for (int32_t k = 0; k < S.n; ++k) {
    const I4 ci = C.get(k, csrc);  // column c = n-1-k
    float corr = 0.f;
    // groups of four: independent window reads, one accumulation chain
    for (; nl - 4 - i >= ci.y; i += 4) {
      E.ensure(i + 3, esrc);
      if constexpr (!SHL) Ls.ensure(i + 3, lsrc);
      const I2 e0 = E.at(i), e1 = E.at(i + 1), e2 = E.at(i + 2), e3 = E.at(i + 3);
      const float l0 = SHL ? __int_as_float(e0.y) : Ls.at(i);
      const float l1 = SHL ? __int_as_float(e1.y) : Ls.at(i + 1);
      const float l2 = SHL ? __int_as_float(e2.y) : Ls.at(i + 2);
      const float l3 = SHL ? __int_as_float(e3.y) : Ls.at(i + 3);
      const float w0 = W[(e0.x & (kFirstTouchBit - 1)) * kWave];
      const float w1 = W[(e1.x & (kFirstTouchBit - 1)) * kWave];
      const float w2 = W[(e2.x & (kFirstTouchBit - 1)) * kWave];
      const float w3 = W[(e3.x & (kFirstTouchBit - 1)) * kWave];
      corr += l0 * w0;
      corr += l1 * w1;
      corr += l2 * w2;
      corr += l3 * w3;
    }
    for (; nl - 1 - i >= ci.y; ++i) {
      const I2 e = E.get(i, esrc);
      const float l = SHL ? __int_as_float(e.y) : Ls.get(i, lsrc);
      corr += l * W[(e.x & (kFirstTouchBit - 1)) * kWave];
    }
    const float w = Ss.get(k, ssrc) - corr;
    if (ci.x >= 0) W[ci.x * kWave] = w;
    out(S.n - 1 - k, ci.z, w);
  }
}

}  // namespace wmpc
