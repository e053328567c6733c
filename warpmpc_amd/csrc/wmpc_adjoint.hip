// OSQP adjoint derivative for MI355X (gfx950) behind include/warpmpc_hip.h.
//
// One launch, one lane per instance, restating warpmpc/jax_osqp/solver.py:
// 1047-1192.  With G = [-A_lower; A_upper], h = [-l_lower; u_upper], the split
// multipliers lambda = [max(-y, 0)_lower; max(y, 0)_upper] and the slacks
// s = G x - h, the adjoint matrix is
//     M = [[(1 + eps) I, U], [Uᵀ, -eps I]],
//     U = [[P, Gᵀ diag(lambda), Aeqᵀ], [G, diag(s), 0], [Aeq, 0, 0]],
// and M0 the same with eps = 0.  Per instance the kernel
//   1. evaluates lambda and s (G x by rows of G) into the work area;
//   2. factors M with the up-looking LDLᵀ of wmpc_qdldl_device.h, its values
//      evaluated slot by slot from (P, A, lambda, s), the forward
//      substitution of rhs = [-dx; -dy_lower; -dy_upper; -d_nu; 0] fused in;
//   3. runs the backward sweep into sol, then `refine` steps
//      sol -= M⁻¹ (M0 sol - rhs), M0 sol by rows of the full symmetric M0;
//   4. maps sol back to dq, dl, du and forms dP, dA.
// Every table the kernel reads is wave-uniform (scalar loads); every
// per-instance array is batch-major [entry][ld].
#include <algorithm>
#include <string>
#include <vector>

#include "wmpc_blob.h"
#include "wmpc_plan.h"
#include "wmpc_stream.h"

namespace wmpc {

// Value sources of an adjoint slot {kind, i0, i1, f bits}:
enum : int32_t {
  kSlotConst = 0,    // f; i0 holds the bits of the perturbed value
  kSlotP = 1,        // p[i0]
  kSlotA = 2,        // a[i0] * f
  kSlotGLambda = 3,  // lambda[i1] * (a[i0] * f)
  kSlotSlack = 4,    // s[i0]
};
// Right-hand-side sources {kind, index} of adjoint row o.
enum : int32_t { kRhsDx = 0, kRhsDy = 1, kRhsEq = 2, kRhsZero = 3 };

struct AdjDev {
  int32_t n = 0, m = 0, ni = 0, half = 0, dim = 0, nnz_p = 0, nnz_a = 0;
  const int4* slot = nullptr;      // [nnz_adj]
  const int32_t* mv_ptr = nullptr; // [dim + 1] rows of the full symmetric M0
  const int2* mv = nullptr;        // {col, slot}
  const int32_t* g_ptr = nullptr;  // [ni + 1] rows of G
  const int4* g = nullptr;         // {col, A value index, sign bits, 0}
  const int2* ineq = nullptr;      // [ni] {constraint index, 1 if upper-bounded}
  const int2* rsrc = nullptr;      // [dim] right-hand-side source
  const int4* con = nullptr;       // [m] {lower, upper, equality position in sol or -1, 0}
  const int2* pent = nullptr;      // [nnz_p] {row, col}
  const int2* aent = nullptr;      // [nnz_a] {row, col}
  // the rows of M0 again, in the sweeps' first-touch order with the value
  // and right-hand-side sources inlined, so the refinement's residual reads
  // two sequential streams: rows {original row, end of its entries, rhs kind,
  // rhs index}, entries {col | kind << 28, i0, i1, f bits}
  int32_t nmv = 0;
  const int4* mrow = nullptr;  // [dim]
  const int4* ment = nullptr;  // [nmv]
};

}  // namespace wmpc

// Tables of the KKT-form adjoint (wmpc_adjoint_kkt_plan_create): the
// converged limit of the refinement above, solved on the forward KKT pattern.
namespace wmpc {
struct AdjKktDev {
  int32_t n = 0, m = 0, nk = 0, nl = 0, ni = 0, half = 0, nnz_p = 0, nnz_a = 0;
  float delta = 0.f;
  const int4* slot = nullptr;       // [nnz_kkt] {kind, index, 0, perturbed offset bits}
  const int32_t* mv_ptr = nullptr;  // [nk + 1] rows of the full symmetric KKT
  const int2* mv = nullptr;         // {col, slot}
  const int32_t* arow_ptr = nullptr;  // [m + 1] rows of A
  const int2* arow = nullptr;         // {col, A value index}
  const int4* con = nullptr;  // [m] {lower position k or -1, upper position k or -1, eq position or -1, 0}
  const int2* pent = nullptr;  // [nnz_p] {row, col}
  const int2* aent = nullptr;  // [nnz_a] {row, col}
  // the rows of the full symmetric KKT again, in the sweeps' first-touch order
  // with each entry's value source inlined, so the residual loops read two
  // sequential streams: rows {original row, end of its entries}, entries
  // {col, kind, index, perturbed offset bits}
  int32_t nmv = 0;
  const int2* mrow = nullptr;  // [nk]
  const int4* ment = nullptr;  // [nmv]
};
}  // namespace wmpc

struct wmpc_adjoint_plan {
  wmpc::AdjDev d;
  wmpc::AdjKktDev k;
  bool kkt_form = false;
  const wmpc_qdldl_plan* adj = nullptr;  // the factored pattern: adjoint or forward KKT
  void* block = nullptr;
};

namespace wmpc {

// Adjoint values of one instance.  PERT: the factored matrix M (eps on both
// diagonals), else M0.  P/A are [nnz][pm_ld] per instance or shared columns.
template <bool PERT>
struct AdjSource {
  const int4* __restrict__ slot;
  const float* __restrict__ p;
  const float* __restrict__ a;
  int64_t pes;   // entry stride of P/A
  uint32_t pl;   // lane offset into P/A (0 when shared)
  CStrided lam, slk;
  __device__ __forceinline__ float operator()(int32_t s, int64_t = 0) const {
    const int4 t = slot[s];
    return from(t.x, t.y, t.z, t.w);
  }
  // Entry {col | kind << 28, i0, i1, f bits} of a row stream.
  __device__ __forceinline__ float entry(const int4 e) const {
    return from((int32_t)((uint32_t)e.x >> 28), e.y, e.z, e.w);
  }
  __device__ __forceinline__ float from(int32_t kind, int32_t i0, int32_t i1, int32_t fb) const {
    switch (kind) {
      case kSlotConst:
        return __int_as_float(PERT ? i0 : fb);
      case kSlotP:
        return (p + (int64_t)i0 * pes)[pl];
      case kSlotA:
        return (a + (int64_t)i0 * pes)[pl] * __int_as_float(fb);
      case kSlotGLambda:
        return lam(i1) * ((a + (int64_t)i0 * pes)[pl] * __int_as_float(fb));
      default:
        return slk(i0);
    }
  }
};

// Right-hand side of adjoint row o (solver.py:1118-1124).
struct AdjRhs {
  const int2* __restrict__ rsrc;
  CStrided dx, dy, y;
  __device__ __forceinline__ float at(int32_t o) const {
    const int2 r = rsrc[o];
    return from(r.x, r.y);
  }
  __device__ __forceinline__ float from(int32_t kind, int32_t idx) const {
    if (kind == kRhsDx) return -dx(idx);
    if (kind == kRhsDy) return -dy(idx);
    if (kind == kRhsEq) {
      const float d = dy(idx);
      return -(y(idx) >= 0.0f ? d : -d);
    }
    return 0.0f;
  }
  __device__ __forceinline__ float operator()(int32_t, int32_t o) const { return at(o); }
};

struct SolSet {
  Strided sol;
  __device__ __forceinline__ void operator()(int32_t, int32_t o, float v) const { sol(o) = v; }
};
struct SolSub {
  Strided sol;
  __device__ __forceinline__ void operator()(int32_t, int32_t o, float v) const {
    sol(o) = sol(o) - v;
  }
};

__global__ __launch_bounds__(kWave) void k_osqp_adjoint(
    QdldlProg P, AdjDev dv, const float* __restrict__ pv, const float* __restrict__ av,
    int64_t pm_ld, const float* __restrict__ l, const float* __restrict__ u,
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ dx,
    const float* __restrict__ dy, float* __restrict__ lx, float* __restrict__ dinv,
    float* __restrict__ stash, float* __restrict__ rbuf, float* __restrict__ sol,
    float* __restrict__ lam, float* __restrict__ slk, float* __restrict__ dq,
    float* __restrict__ dl, float* __restrict__ du, float* __restrict__ dp,
    float* __restrict__ da, int32_t refine, int64_t batch, int64_t ld, int32_t ipw) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  // every lane takes part in the wave-shared staging of the sweeps, lanes
  // past the batch on the last instance
  int64_t b0;
  const int64_t b = lane_instance(ipw, batch, &b0);
  float* mywin = lds + (lane & (P.lcols - 1));
  const int64_t pes = pm_ld ? pm_ld : 1;
  const uint32_t pl = pm_ld ? (uint32_t)b : 0u;
  const CStrided X{x, ld, b}, Y{y, ld, b}, Lo{l, ld, b}, Up{u, ld, b};
  const Strided LAM{lam, ld, b}, SLK{slk, ld, b}, SOL{sol, ld, b}, RB{rbuf, ld, b};
  // 1. multipliers and slacks of the inequality rows (solver.py:1082-1097)
  for (int32_t i = 0; i < dv.ni; ++i) {
    const int2 q = dv.ineq[i];
    const float yv = Y(q.x);
    float s = 0.0f;
    for (int32_t e = dv.g_ptr[i]; e < dv.g_ptr[i + 1]; ++e) {
      const int4 t = dv.g[e];
      const float gv = (av + (int64_t)t.y * pes)[pl] * __int_as_float(t.z);
      s += gv * X(t.x);
    }
    LAM(i) = q.y ? nan_max(yv, 0.0f) : -nan_min(yv, 0.0f);
    SLK(i) = s - (q.y ? Up(q.x) : -Lo(q.x));
  }
  const CStrided CLAM{lam, ld, b}, CSLK{slk, ld, b};
  const AdjSource<true> msrc{dv.slot, pv, av, pes, pl, CLAM, CSLK};
  const AdjSource<false> m0{dv.slot, pv, av, pes, pl, CLAM, CSLK};
  const AdjRhs rhs{dv.rsrc, CStrided{dx, ld, b}, CStrided{dy, ld, b}, Y};
  // 2. factor M with the forward substitution of rhs fused in
  factor_rows<true>(P, msrc, 0, Strided{lx, ld, b}, Strided{nullptr, 0, 0}, Strided{dinv, ld, b},
                    mywin, rhs, Strided{stash, ld, b});
  // 3. backward sweep, then the refinement against M0 (solver.py:1127-1139)
  const SweepProg SP = sweep_prog(P);
  bwd_cols<false>(SP, lx, ld, b, stash, ld, b, lds, lane, SolSet{SOL});
  for (int32_t it = 0; it < refine; ++it) {
    // residual M0 sol - rhs, staged in the sweep's first-touch order
    TableStream<int4> rows(dv.mrow, dv.dim, 0), ents(dv.ment, dv.nmv, 0);
    int32_t e = 0;
    for (int32_t i = 0; i < dv.dim; ++i) {
      const int4 r = rows[i];
      float s = 0.0f;
      // four entries' operands loaded before they are summed in order
      for (; e + 4 <= r.y; e += 4) {
        int4 t[4];
        float v[4], x[4];
        ents.get4(e, t[0], t[1], t[2], t[3]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v[j] = m0.entry(t[j]);
          x[j] = SOL(t[j].x & 0x0FFFFFFF);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) s += v[j] * x[j];
      }
      for (; e < r.y; ++e) {
        const int4 t = ents[e];
        s += m0.entry(t) * SOL(t.x & 0x0FFFFFFF);
      }
      RB(i) = s - rhs.from(r.z, r.w);
    }
    fwd_cols<false>(SP, lx, dinv, ld, b, rbuf, stash, ld, b, lds, lane);
    bwd_cols<false>(SP, lx, ld, b, stash, ld, b, lds, lane, SolSub{SOL});
  }
  // 4. outputs (solver.py:1141-1192); only real lanes store
  if (b0 >= batch) return;
  const Strided DQ{dq, ld, b}, DL{dl, ld, b}, DU{du, ld, b};
  const int32_t n = dv.n, half = dv.half;
  for (int32_t i = 0; i < n; ++i) DQ(i) = SOL(half + i);
  for (int32_t i = 0; i < dv.m; ++i) {
    const int4 t = dv.con[i];
    const float yv = Y(i);
    const float yu = nan_max(yv, 0.0f), yl = -nan_min(yv, 0.0f);
    float rl = t.x >= 0 ? -SOL(t.x) : 0.0f;
    float ru = t.y >= 0 ? SOL(t.y) : 0.0f;
    if (t.z >= 0) {
      const float es = SOL(t.z);
      const bool pos = yv >= 0.0f;
      ru = pos ? es / yv : 0.0f;
      rl = pos ? 0.0f : -es / yv;
    }
    DL(i) = -(rl * yl);
    DU(i) = -(ru * yu);
  }
  if (dp) {
    const Strided DP{dp, ld, b};
    for (int32_t k = 0; k < dv.nnz_p; ++k) {
      const int2 t = dv.pent[k];
      DP(k) = 0.5f * (SOL(half + t.x) * X(t.y) + SOL(half + t.y) * X(t.x));
    }
  }
  if (da) {
    // (y_u - y_l) = y and (ryu - ryl) = -du - dl, exactly
    const Strided DA{da, ld, b};
    for (int32_t k = 0; k < dv.nnz_a; ++k) {
      const int2 t = dv.aent[k];
      DA(k) = Y(t.x) * SOL(half + t.y) + (-DU(t.x) - DL(t.x)) * X(t.y);
    }
  }
}

// ---------------------------------------------------------------------------
// KKT form of the adjoint: implicit differentiation of the active-set KKT
// system.  With U nonsingular, the limit of the refinement above solves
// M0 [a; b] = [r; 0], i.e. a = 0 and U b = r.  Writing c_i = ±lambda b_lambda
// for the inequality row of A row i that has a positive multiplier (sign: +
// upper, - lower) and c_i = b_nu for an equality row, U b = r becomes the
// symmetric system on the forward KKT pattern
//     [[P, Aᵀ], [A, diag(d)]] [b_x; c] = [-dx; rc],
//     d_i = min(s, 0)/lambda (inequality with a positive multiplier),
//           0 (equality), -1e30 (no positive multiplier: the row decouples,
//           c_i = 0),
//     rc_i = -dy_i (upper), dy_i (lower), -d_nu_i (equality), 0.
// Where the rows with a positive multiplier are feasible (s <= 0, as at a
// converged solution) this is exactly the reference's limit.  The kernel
// equilibrates the system per instance (Ruiz, 2 passes) before factoring it
// in float32: without that, element growth on an equality-heavy KKT (the
// quadrotor MPC) makes the refinement diverge.  An iterate of a
// few ADMM iterations can violate such a row (s > 0); the slack is then taken
// as 0, the value at the converged solution: the reference's U is nearly
// singular at such points, the active-set system is not.  d <= 0 always, so
// the matrix factored with sigma on P's diagonal and d - delta on the
// constraint diagonal is quasi-definite; the kernel refines against the
// exact matrix, and the outputs follow solver.py:1141-1192 with
// b_lambda = ±c/lambda.
enum : int32_t { kKktP = 0, kKktA = 1, kKktDiag = 2 };
constexpr float kDecoupled = -1e30f;

// Values of the KKT-form matrix.  PERT: the factored matrix (sigma on P's
// diagonal, -delta on the constraint diagonal), else the exact one.  SCALED:
// equilibrated, K_s = S K S with S = diag(sc) (slot.z packs the entry's row
// and column); a decoupled row's -1e30 stays unscaled.
template <bool PERT, bool SCALED>
struct AdjKktSource {
  const int4* __restrict__ slot;
  const float* __restrict__ p;
  const float* __restrict__ a;
  int64_t pes;
  uint32_t pl;
  CStrided dg, sc;
  float delta;
  __device__ __forceinline__ float operator()(int32_t s, int64_t = 0) const {
    const int4 t = slot[s];
    float f = 1.0f;
    if constexpr (SCALED) f = sc(t.z & 0xffff) * sc((uint32_t)t.z >> 16);
    return from(t.x, t.y, t.w, f);
  }
  // Entry {col, kind, index, perturbed offset bits} of row o (the scale
  // product is commutative, so sc(o) sc(col) equals the upper entry's).
  __device__ __forceinline__ float entry(const int4 e, int32_t o) const {
    float f = 1.0f;
    if constexpr (SCALED) f = sc(o) * sc(e.x);
    return from(e.y, e.z, e.w, f);
  }
  __device__ __forceinline__ float from(int32_t kind, int32_t idx, int32_t pert, float f) const {
    if (kind == kKktP) {
      const float v = (p + (int64_t)idx * pes)[pl];
      return (PERT ? v + __int_as_float(pert) : v) * f;
    }
    if (kind == kKktA) return (a + (int64_t)idx * pes)[pl] * f;
    const float d = dg(idx);
    if (d <= kDecoupled) return d;
    return PERT ? d * f - delta : d * f;
  }
};

struct AdjKktRhs {
  int32_t n;
  CStrided dx, rc, sc;
  __device__ __forceinline__ float at(int32_t o) const { return (o < n ? -dx(o) : rc(o - n)) * sc(o); }
  __device__ __forceinline__ float operator()(int32_t, int32_t o) const { return at(o); }
};

// Ruiz passes of the symmetric equilibration
constexpr int kRuizPasses = 2;

__global__ __launch_bounds__(kWave) void k_osqp_adjoint_kkt(
    QdldlProg P, AdjKktDev dv, const float* __restrict__ pv, const float* __restrict__ av,
    int64_t pm_ld, const float* __restrict__ l, const float* __restrict__ u,
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ dx,
    const float* __restrict__ dy, float* __restrict__ lx, float* __restrict__ dinv,
    float* __restrict__ stash, float* __restrict__ rbuf, float* __restrict__ w,
    float* __restrict__ dg, float* __restrict__ rc, float* __restrict__ axv,
    float* __restrict__ scv, float* __restrict__ sol, float* __restrict__ dq, float* __restrict__ dl,
    float* __restrict__ du, float* __restrict__ dp, float* __restrict__ da, int32_t refine,
    int64_t batch, int64_t ld, int32_t ipw) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  int64_t b0;
  const int64_t b = lane_instance(ipw, batch, &b0);
  float* mywin = lds + (lane & (P.lcols - 1));
  const int64_t pes = pm_ld ? pm_ld : 1;
  const uint32_t pl = pm_ld ? (uint32_t)b : 0u;
  const CStrided X{x, ld, b}, Y{y, ld, b}, Lo{l, ld, b}, Up{u, ld, b}, DY{dy, ld, b};
  const Strided DG{dg, ld, b}, RC{rc, ld, b}, AX{axv, ld, b}, W{w, ld, b}, RB{rbuf, ld, b};
  // 1. per A row: A x, the coupling diagonal and the right-hand side (tables
  // as streams; a row's operands loaded four at a time, summed in order)
  TableStream<int32_t> aends(dv.arow_ptr + 1, dv.m, 0);
  TableStream<int2> aents(dv.arow, dv.arow_ptr[dv.m], 0);
  TableStream<int4> cons(dv.con, dv.m, 0);
  int32_t ae = dv.arow_ptr[0];
  for (int32_t i = 0; i < dv.m; ++i) {
    float ax = 0.0f;
    const int32_t aend = aends[i];
    for (; ae + 4 <= aend; ae += 4) {
      int2 ta[4];
      float va[4], xa[4];
      aents.get4(ae, ta[0], ta[1], ta[2], ta[3]);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        va[j] = (av + (int64_t)ta[j].y * pes)[pl];
        xa[j] = X(ta[j].x);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) ax += va[j] * xa[j];
    }
    for (; ae < aend; ++ae) {
      const int2 ta = aents[ae];
      ax += (av + (int64_t)ta.y * pes)[pl] * X(ta.x);
    }
    const int4 t = cons[i];
    const float yv = Y(i), dyv = DY(i);
    const float lu = nan_max(yv, 0.0f), ll = -nan_min(yv, 0.0f);
    float d = kDecoupled, r = 0.0f;
    if (t.z >= 0) {
      d = 0.0f;
      r = -(yv >= 0.0f ? dyv : -dyv);
    } else if (t.y >= 0 && lu > 0.0f) {
      d = fminf((ax - Up(i)) / lu, 0.0f);
      r = -dyv;
    } else if (t.x >= 0 && ll > 0.0f) {
      d = fminf((Lo(i) - ax) / ll, 0.0f);
      r = dyv;
    }
    if (yv != yv) d = r = yv;
    DG(i) = d;
    RC(i) = r;
    AX(i) = ax;
  }
  const CStrided CDG{dg, ld, b}, CSC{scv, ld, b};
  const Strided SC{scv, ld, b};
  // 2. symmetric equilibration (Ruiz): sc_o <- sc_o / sqrt(max_j |sc_o K_oj sc_j|),
  // a decoupled row's diagonal left out; the new scales are staged in rbuf
  {
    const AdjKktSource<false, false> raw{dv.slot, pv, av, pes, pl, CDG, CSC, dv.delta};
    for (int32_t o = 0; o < dv.nk; ++o) SC(o) = 1.0f;
    for (int32_t pass = 0; pass < kRuizPasses; ++pass) {
      // rows in first-touch order (each row's scale depends only on the
      // previous pass)
      TableStream<int2> rows(dv.mrow, dv.nk, 0);
      TableStream<int4> ents(dv.ment, dv.nmv, 0);
      int32_t e = 0;
      for (int32_t i = 0; i < dv.nk; ++i) {
        const int2 r = rows[i];
        const int32_t o = r.x;
        float mx = 0.0f;
        for (; e < r.y; ++e) {
          const int4 t = ents[e];
          const float v = raw.entry(t, o);
          if (v > kDecoupled) mx = fmaxf(mx, fabsf(v) * SC(t.x));
        }
        mx *= SC(o);
        RB(o) = mx > 0.0f ? SC(o) / sqrtf(mx) : SC(o);
      }
      for (int32_t o = 0; o < dv.nk; ++o) SC(o) = RB(o);
    }
  }
  const AdjKktSource<true, true> msrc{dv.slot, pv, av, pes, pl, CDG, CSC, dv.delta};
  const AdjKktSource<false, true> m0{dv.slot, pv, av, pes, pl, CDG, CSC, dv.delta};
  const AdjKktRhs rhs{dv.n, CStrided{dx, ld, b}, CStrided{rc, ld, b}, CSC};
  // 3. factor the equilibrated matrix with the forward substitution fused
  // in, backward sweep (W holds the equilibrated solution S⁻¹ w)
  factor_rows<true>(P, msrc, 0, Strided{lx, ld, b}, Strided{nullptr, 0, 0}, Strided{dinv, ld, b},
                    mywin, rhs, Strided{stash, ld, b});
  const SweepProg SP = sweep_prog(P);
  bwd_cols<false>(SP, lx, ld, b, stash, ld, b, lds, lane, SolSet{W});
  // 4. refinement against the exact (equilibrated) matrix
  for (int32_t it = 0; it < refine; ++it) {
    TableStream<int2> rows(dv.mrow, dv.nk, 0);
    TableStream<int4> ents(dv.ment, dv.nmv, 0);
    int32_t e = 0;
    for (int32_t i = 0; i < dv.nk; ++i) {
      const int2 r = rows[i];
      float s = 0.0f;
      // four entries' operands loaded before they are summed in order
      for (; e + 4 <= r.y; e += 4) {
        int4 t[4];
        float v[4], w[4];
        ents.get4(e, t[0], t[1], t[2], t[3]);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          v[j] = m0.entry(t[j], r.x);
          w[j] = W(t[j].x);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) s += v[j] * w[j];
      }
      for (; e < r.y; ++e) {
        const int4 t = ents[e];
        s += m0.entry(t, r.x) * W(t.x);
      }
      RB(i) = s - rhs.at(r.x);
    }
    fwd_cols<false>(SP, lx, dinv, ld, b, rbuf, stash, ld, b, lds, lane);
    bwd_cols<false>(SP, lx, ld, b, stash, ld, b, lds, lane, SolSub{W});
  }
  // the solution w = S W, in place
  for (int32_t o = 0; o < dv.nk; ++o) W(o) = W(o) * SC(o);
  // 5. outputs; only real lanes store
  if (b0 >= batch) return;
  const Strided DQ{dq, ld, b}, DL{dl, ld, b}, DU{du, ld, b};
  const int32_t n = dv.n, half = dv.half;
  for (int32_t j = 0; j < n; ++j) DQ(j) = W(j);
  for (int32_t i = 0; i < dv.m; ++i) {
    const int4 t = dv.con[i];
    const float yv = Y(i), c = W(n + i);
    const float yu = nan_max(yv, 0.0f), yl = -nan_min(yv, 0.0f);
    float rl = 0.0f, ru = 0.0f;
    if (t.z >= 0) {
      const bool pos = yv >= 0.0f;
      ru = pos ? c / yv : 0.0f;
      rl = pos ? 0.0f : -c / yv;
    } else {
      // b_lambda of a row with a positive multiplier is ±c/lambda; a row
      // without one has lambda = 0 and contributes 0 * b_lambda
      if (t.y >= 0 && yu > 0.0f) ru = c / yu;
      if (t.x >= 0 && yl > 0.0f) rl = c / yl;
    }
    DL(i) = -(rl * yl);
    DU(i) = -(ru * yu);
  }
  if (sol) {
    // the M0 solution [a; b] = [0; b_x, b_lambda, b_nu] of the reference
    const Strided SOL{sol, ld, b};
    for (int32_t o = 0; o < half; ++o) SOL(o) = 0.0f;
    for (int32_t j = 0; j < n; ++j) SOL(half + j) = W(j);
    const int32_t base = half + n;
    for (int32_t i = 0; i < dv.m; ++i) {
      const int4 t = dv.con[i];
      const float yv = Y(i), c = W(n + i), dyv = DY(i);
      if (t.z >= 0) {
        SOL(base + t.z) = c;
        continue;
      }
      if (t.x < 0 && t.y < 0) continue;
      float abx = 0.0f;  // A_i b_x
      for (int32_t e = dv.arow_ptr[i]; e < dv.arow_ptr[i + 1]; ++e) {
        const int2 a2 = dv.arow[e];
        abx += (av + (int64_t)a2.y * pes)[pl] * W(a2.x);
      }
      const float ax = AX(i);
      // a row coupled through c gives b_lambda = ±c/lambda; a decoupled one
      // (no positive multiplier, or one so small that s/lambda passed -1e30)
      // gets it from its own equation, G_i b_x + s b_lambda = -dy
      const bool coupled = DG(i) > kDecoupled;
      if (t.x >= 0) {  // lower row: G = -A_i, s = l - A x
        const float lam = -nan_min(yv, 0.0f);
        SOL(base + t.x) = (coupled && lam > 0.0f) ? -c / lam : (-dyv + abx) / (Lo(i) - ax);
      }
      if (t.y >= 0) {  // upper row: G = A_i, s = A x - u
        const float lam = nan_max(yv, 0.0f);
        SOL(base + t.y) = (coupled && lam > 0.0f) ? c / lam : (-dyv - abx) / (ax - Up(i));
      }
    }
  }
  // dP and dA, 8 entries per step with their loads first (only real lanes
  // run here, so the entry tables are read directly, not as streams)
  if (dp) {
    const Strided DP{dp, ld, b};
    int32_t k = 0;
    for (; k + 8 <= dv.nnz_p; k += 8) {
      int2 t[8];
      float wx[8], xy[8], wy[8], xx[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) t[j] = dv.pent[k + j];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        wx[j] = W(t[j].x);
        xy[j] = X(t[j].y);
        wy[j] = W(t[j].y);
        xx[j] = X(t[j].x);
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) DP(k + j) = 0.5f * (wx[j] * xy[j] + wy[j] * xx[j]);
    }
    for (; k < dv.nnz_p; ++k) {
      const int2 t = dv.pent[k];
      DP(k) = 0.5f * (W(t.x) * X(t.y) + W(t.y) * X(t.x));
    }
  }
  if (da) {
    const Strided DA{da, ld, b};
    int32_t k = 0;
    for (; k + 8 <= dv.nnz_a; k += 8) {
      int2 t[8];
      float yx[8], wy[8], ux[8], lxv[8], xy[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) t[j] = dv.aent[k + j];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        yx[j] = Y(t[j].x);
        wy[j] = W(t[j].y);
        ux[j] = DU(t[j].x);
        lxv[j] = DL(t[j].x);
        xy[j] = X(t[j].y);
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) DA(k + j) = yx[j] * wy[j] + (-ux[j] - lxv[j]) * xy[j];
    }
    for (; k < dv.nnz_a; ++k) {
      const int2 t = dv.aent[k];
      DA(k) = Y(t.x) * W(t.y) + (-DU(t.x) - DL(t.x)) * X(t.y);
    }
  }
}

static size_t adjoint_lds(const QdldlProg& P) {
  return (size_t)std::max(factor_lds_floats(P, true), sweep_lds_floats(sweep_prog(P), false)) *
         sizeof(float);
}

}  // namespace wmpc

using namespace wmpc;

extern "C" {

int wmpc_adjoint_plan_create(const wmpc_adjoint_desc* ds, const wmpc_qdldl_plan* adj,
                             wmpc_adjoint_plan** out) {
  clear_error();
  WMPC_REQUIRE(ds && adj && out, "null argument");
  *out = nullptr;
  WMPC_REQUIRE(adj->block, "host-only adjoint plan: create it with wmpc_qdldl_plan_create");
  const int32_t n = ds->n, m = ds->m, nl = ds->n_l, nu = ds->n_u, ne = ds->n_eq;
  WMPC_REQUIRE(n > 0 && m >= 0 && nl >= 0 && nu >= 0 && ne >= 0 && nl <= m && nu <= m && ne <= m,
               "bad dimensions");
  const int32_t ni = nl + nu, half = n + ni + ne, dim = 2 * half, nz = ds->nnz_adj;
  WMPC_REQUIRE(adj->host.n == dim, "adjoint QDLDL plan dimension must be 2*(n + n_l + n_u + n_eq)");
  WMPC_REQUIRE(adj->host.nnz_a == nz, "adjoint nonzero count differs from the QDLDL plan");
  WMPC_REQUIRE(ds->nnz_p >= 0 && ds->nnz_a >= 0 && ds->n_p_full >= 0 && ds->n_g >= 0 && ds->n_aeq >= 0,
               "bad counts");
  WMPC_REQUIRE(ds->adj_indptr && ds->adj_indices && ds->adj_data && ds->diag_first_pos &&
                   ds->diag_second_pos && (nl == 0 || ds->lower_indices) &&
                   (nu == 0 || ds->upper_indices) && (ne == 0 || ds->eq_indices) &&
                   (ds->nnz_p == 0 || (ds->p_upper_rows && ds->p_upper_cols)) &&
                   (ds->n_p_full == 0 || (ds->p_full_value_map && ds->p_block_pos)) &&
                   (ds->n_g == 0 || (ds->g_rows && ds->g_cols && ds->g_a_value_map && ds->g_sign &&
                                     ds->g_block_pos && ds->glambda_block_pos)) &&
                   (ds->n_aeq == 0 || (ds->aeq_a_value_map && ds->aeq_block_pos && ds->aeq_t_block_pos)) &&
                   (ni == 0 || ds->slack_diag_pos) && (ds->nnz_a == 0 || (ds->a_rows && ds->a_cols)),
               "descriptor has a null array");
  // CSC structure of the upper triangle
  WMPC_REQUIRE(ds->adj_indptr[0] == 0 && ds->adj_indptr[dim] == nz, "adj_indptr does not span nnz_adj");
  std::vector<int32_t> rows(nz), cols(nz);
  for (int32_t c = 0; c < dim; ++c) {
    WMPC_REQUIRE(ds->adj_indptr[c] <= ds->adj_indptr[c + 1], "adj_indptr not monotone");
    for (int32_t k = ds->adj_indptr[c]; k < ds->adj_indptr[c + 1]; ++k) {
      const int32_t r = ds->adj_indices[k];
      WMPC_REQUIRE(r >= 0 && r <= c, "adjoint entry outside the upper triangle");
      rows[k] = r;
      cols[k] = c;
    }
  }
  // slot sources
  std::vector<int32_t> slot(4 * (size_t)nz, 0);
  std::vector<char> set(nz, 0);
  for (int32_t k = 0; k < nz; ++k) {
    slot[4 * k] = kSlotConst;
    slot[4 * k + 1] = float_bits(ds->adj_data[k]);
    slot[4 * k + 3] = float_bits(ds->adj_data[k]);
  }
  std::string err;
  auto claim = [&](int32_t pos, int32_t kind, int32_t i0, int32_t i1, double f) {
    if (pos < 0 || pos >= nz || set[pos]) {
      err = "adjoint block position out of range or repeated";
      return;
    }
    set[pos] = 1;
    slot[4 * pos] = kind;
    slot[4 * pos + 1] = i0;
    slot[4 * pos + 2] = i1;
    slot[4 * pos + 3] = float_bits(f);
  };
  for (int32_t i = 0; i < half; ++i) {
    claim(ds->diag_first_pos[i], kSlotConst, float_bits(1.0 + ds->eps), 0, 1.0);
    claim(ds->diag_second_pos[i], kSlotConst, float_bits(-ds->eps), 0, 0.0);
  }
  for (int32_t k = 0; k < ds->n_p_full; ++k) {
    const int32_t v = ds->p_full_value_map[k];
    WMPC_REQUIRE(v >= 0 && v < ds->nnz_p, "p_full_value_map out of range");
    claim(ds->p_block_pos[k], kSlotP, v, 0, 1.0);
  }
  for (int32_t k = 0; k < ds->n_g; ++k) {
    const int32_t v = ds->g_a_value_map[k], r = ds->g_rows[k], c = ds->g_cols[k];
    WMPC_REQUIRE(v >= 0 && v < ds->nnz_a && r >= 0 && r < ni && c >= 0 && c < n, "G entry out of range");
    claim(ds->g_block_pos[k], kSlotA, v, 0, ds->g_sign[k]);
    claim(ds->glambda_block_pos[k], kSlotGLambda, v, r, ds->g_sign[k]);
  }
  for (int32_t k = 0; k < ds->n_aeq; ++k) {
    const int32_t v = ds->aeq_a_value_map[k];
    WMPC_REQUIRE(v >= 0 && v < ds->nnz_a, "aeq_a_value_map out of range");
    claim(ds->aeq_block_pos[k], kSlotA, v, 0, 1.0);
    claim(ds->aeq_t_block_pos[k], kSlotA, v, 0, 1.0);
  }
  for (int32_t i = 0; i < ni; ++i) claim(ds->slack_diag_pos[i], kSlotSlack, i, 0, 0.0);
  WMPC_REQUIRE(err.empty(), err.c_str());
  // rows of the full symmetric M0: the row's upper entries (columns >= o,
  // ascending), then the mirrored ones (rows < o, ascending) -- the order in
  // which the reference's scatter matvec accumulates them
  std::vector<int32_t> mv_ptr(dim + 1, 0), mv;
  {
    std::vector<std::vector<int32_t>> up(dim), lo(dim);
    for (int32_t k = 0; k < nz; ++k) {
      up[rows[k]].push_back(k);
      if (rows[k] != cols[k]) lo[cols[k]].push_back(k);
    }
    for (int32_t o = 0; o < dim; ++o) {
      for (int32_t k : up[o]) mv.insert(mv.end(), {cols[k], k});
      for (int32_t k : lo[o]) mv.insert(mv.end(), {rows[k], k});
      mv_ptr[o + 1] = (int32_t)(mv.size() / 2);
    }
  }
  // rows of G (entries within a row by ascending column, as the scatter of
  // solver.py:1089-1092 accumulates them)
  std::vector<int32_t> g_ptr(ni + 1, 0), g;
  {
    std::vector<std::vector<int32_t>> byrow(ni);
    for (int32_t k = 0; k < ds->n_g; ++k) byrow[ds->g_rows[k]].push_back(k);
    for (int32_t i = 0; i < ni; ++i) {
      std::stable_sort(byrow[i].begin(), byrow[i].end(),
                       [&](int32_t a, int32_t b) { return ds->g_cols[a] < ds->g_cols[b]; });
      for (int32_t k : byrow[i])
        g.insert(g.end(), {ds->g_cols[k], ds->g_a_value_map[k], float_bits(ds->g_sign[k]), 0});
      g_ptr[i + 1] = (int32_t)(g.size() / 4);
    }
  }
  std::vector<int32_t> ineq(2 * (size_t)ni), rsrc(2 * (size_t)dim, 0), con(4 * (size_t)m, -1);
  for (int32_t i = 0; i < m; ++i) con[4 * i + 3] = 0;
  for (int32_t o = 0; o < dim; ++o) rsrc[2 * o] = kRhsZero;
  for (int32_t i = 0; i < n; ++i) {
    rsrc[2 * i] = kRhsDx;
    rsrc[2 * i + 1] = i;
  }
  for (int32_t i = 0; i < ni; ++i) {
    const bool upper = i >= nl;
    const int32_t c = upper ? ds->upper_indices[i - nl] : ds->lower_indices[i];
    WMPC_REQUIRE(c >= 0 && c < m, "lower/upper index out of range");
    ineq[2 * i] = c;
    ineq[2 * i + 1] = upper ? 1 : 0;
    rsrc[2 * (n + i)] = kRhsDy;
    rsrc[2 * (n + i) + 1] = c;
    con[4 * c + (upper ? 1 : 0)] = half + n + i;
  }
  for (int32_t i = 0; i < ne; ++i) {
    const int32_t c = ds->eq_indices[i];
    WMPC_REQUIRE(c >= 0 && c < m, "eq index out of range");
    WMPC_REQUIRE(con[4 * c] < 0 && con[4 * c + 1] < 0, "an equality row is also an inequality row");
    rsrc[2 * (n + ni + i)] = kRhsEq;
    rsrc[2 * (n + ni + i) + 1] = c;
    con[4 * c + 2] = half + n + ni + i;
  }
  std::vector<int32_t> pent(2 * (size_t)ds->nnz_p), aent(2 * (size_t)ds->nnz_a);
  for (int32_t k = 0; k < ds->nnz_p; ++k) {
    WMPC_REQUIRE(ds->p_upper_rows[k] >= 0 && ds->p_upper_rows[k] < n && ds->p_upper_cols[k] >= 0 &&
                     ds->p_upper_cols[k] < n,
                 "p_upper_rows/cols out of range");
    pent[2 * k] = ds->p_upper_rows[k];
    pent[2 * k + 1] = ds->p_upper_cols[k];
  }
  for (int32_t k = 0; k < ds->nnz_a; ++k) {
    WMPC_REQUIRE(ds->a_rows[k] >= 0 && ds->a_rows[k] < m && ds->a_cols[k] >= 0 && ds->a_cols[k] < n,
                 "a_rows/a_cols out of range");
    aent[2 * k] = ds->a_rows[k];
    aent[2 * k + 1] = ds->a_cols[k];
  }
  Blob blob;
  // the rows of M0 in the sweeps' first-touch order, sources inlined
  const std::vector<int32_t>& rperm = adj->host.rperm;
  WMPC_REQUIRE((int32_t)rperm.size() == dim, "adjoint plan has no first-touch order");
  WMPC_REQUIRE(dim < (1 << 28), "the adjoint row stream packs columns in 28 bits");
  std::vector<int32_t> mrow(4 * (size_t)dim), ment;
  ment.reserve(2 * mv.size());
  for (int32_t i = 0; i < dim; ++i) {
    const int32_t o = rperm[i];
    WMPC_REQUIRE(o >= 0 && o < dim, "first-touch order out of range");
    for (int32_t e = mv_ptr[o]; e < mv_ptr[o + 1]; ++e) {
      const int32_t k = mv[2 * (size_t)e + 1];
      const int32_t kind = slot[4 * (size_t)k];
      WMPC_REQUIRE(kind >= 0 && kind < 8, "slot kind out of range");
      ment.insert(ment.end(), {mv[2 * (size_t)e] | (kind << 28), slot[4 * (size_t)k + 1],
                               slot[4 * (size_t)k + 2], slot[4 * (size_t)k + 3]});
    }
    mrow[4 * (size_t)i] = o;
    mrow[4 * (size_t)i + 1] = (int32_t)(ment.size() / 4);
    mrow[4 * (size_t)i + 2] = rsrc[2 * (size_t)o];
    mrow[4 * (size_t)i + 3] = rsrc[2 * (size_t)o + 1];
  }
  const size_t t_slot = blob.add(slot), t_mvp = blob.add(mv_ptr), t_mv = blob.add(mv),
               t_gp = blob.add(g_ptr), t_g = blob.add(g), t_in = blob.add(ineq), t_rs = blob.add(rsrc),
               t_con = blob.add(con), t_pe = blob.add(pent), t_ae = blob.add(aent), t_mr = blob.add(mrow),
               t_me = blob.add(ment);
  auto* plan = new wmpc_adjoint_plan();
  if (int rc = blob.upload(&plan->block, "adjoint plan")) {
    delete plan;
    return rc;
  }
  AdjDev& d = plan->d;
  d.n = n;
  d.m = m;
  d.ni = ni;
  d.half = half;
  d.dim = dim;
  d.nnz_p = ds->nnz_p;
  d.nnz_a = ds->nnz_a;
  d.slot = blob.at<int4>(t_slot);
  d.mv_ptr = blob.at<int32_t>(t_mvp);
  d.mv = blob.at<int2>(t_mv);
  d.g_ptr = blob.at<int32_t>(t_gp);
  d.g = blob.at<int4>(t_g);
  d.ineq = blob.at<int2>(t_in);
  d.rsrc = blob.at<int2>(t_rs);
  d.con = blob.at<int4>(t_con);
  d.pent = blob.at<int2>(t_pe);
  d.aent = blob.at<int2>(t_ae);
  d.nmv = (int32_t)(ment.size() / 4);
  d.mrow = blob.at<int4>(t_mr);
  d.ment = blob.at<int4>(t_me);
  plan->adj = adj;
  *out = plan;
  return WMPC_OK;
}

int wmpc_adjoint_plan_destroy(wmpc_adjoint_plan* plan) {
  clear_error();
  if (!plan) return WMPC_OK;
  if (plan->block) WMPC_HIP_CHECK(hipFree(plan->block));
  delete plan;
  return WMPC_OK;
}

int wmpc_adjoint_kkt_plan_create(const wmpc_adjoint_kkt_desc* ds, const wmpc_qdldl_plan* kkt,
                                 wmpc_adjoint_plan** out) {
  clear_error();
  WMPC_REQUIRE(ds && kkt && out, "null argument");
  *out = nullptr;
  WMPC_REQUIRE(kkt->block, "host-only KKT plan: create it with wmpc_qdldl_plan_create");
  const int32_t n = ds->n, m = ds->m, nl = ds->n_l, nu = ds->n_u, ne = ds->n_eq;
  const int32_t nk = n + m, nz = ds->nnz_kkt;
  WMPC_REQUIRE(n > 0 && m >= 0 && nl >= 0 && nu >= 0 && ne >= 0 && nl <= m && nu <= m && ne <= m,
               "bad dimensions");
  WMPC_REQUIRE(kkt->host.n == nk, "KKT plan dimension must be n + m");
  WMPC_REQUIRE(nk <= 65535, "the KKT form packs row and column in 16 bits: n + m must be below 65536");
  WMPC_REQUIRE(kkt->host.nnz_a == nz, "nnz_kkt differs from the KKT plan");
  WMPC_REQUIRE(ds->nnz_p >= 0 && ds->nnz_a >= 0 && nz >= 0, "bad counts");
  WMPC_REQUIRE(ds->delta > 0.0, "delta must be positive");
  WMPC_REQUIRE((nz == 0 || (ds->kkt_rows && ds->kkt_cols)) &&
                   (ds->nnz_p == 0 || (ds->p_to_kkt && ds->p_diag_sigma && ds->p_upper_rows &&
                                       ds->p_upper_cols)) &&
                   (ds->nnz_a == 0 || (ds->a_to_kkt && ds->a_rows && ds->a_cols)) &&
                   (nl == 0 || ds->lower_indices) && (nu == 0 || ds->upper_indices) &&
                   (ne == 0 || ds->eq_indices),
               "descriptor has a null array");
  // value sources of the KKT slots: P, A, or the constraint diagonal
  std::vector<int32_t> slot(4 * (size_t)nz, -1);
  std::vector<char> set(nz, 0);
  for (int32_t k = 0; k < ds->nnz_p; ++k) {
    const int32_t s = ds->p_to_kkt[k];
    WMPC_REQUIRE(s >= 0 && s < nz && !set[s], "p_to_kkt out of range or repeated");
    WMPC_REQUIRE(ds->p_upper_rows[k] >= 0 && ds->p_upper_rows[k] < n && ds->p_upper_cols[k] >= 0 &&
                     ds->p_upper_cols[k] < n,
                 "p_upper_rows/cols out of range");
    set[s] = 1;
    slot[4 * s] = kKktP;
    slot[4 * s + 1] = k;
    slot[4 * s + 3] = float_bits(ds->p_diag_sigma[k]);
  }
  for (int32_t k = 0; k < ds->nnz_a; ++k) {
    const int32_t s = ds->a_to_kkt[k];
    WMPC_REQUIRE(s >= 0 && s < nz && !set[s], "a_to_kkt out of range or repeated");
    set[s] = 1;
    slot[4 * s] = kKktA;
    slot[4 * s + 1] = k;
    slot[4 * s + 3] = 0;
  }
  for (int32_t k = 0; k < nz; ++k) {
    const int32_t r = ds->kkt_rows[k], c = ds->kkt_cols[k];
    WMPC_REQUIRE(r >= 0 && c >= 0 && r <= c && c < nk, "KKT entry outside the upper triangle");
    slot[4 * k + 2] = r | (c << 16);  // row and column of the entry, for the equilibration
    if (set[k]) continue;
    WMPC_REQUIRE(r == c && r >= n, "a KKT slot outside P and A must be a constraint diagonal");
    set[k] = 1;
    slot[4 * k] = kKktDiag;
    slot[4 * k + 1] = r - n;
    slot[4 * k + 3] = 0;
  }
  // rows of the full symmetric KKT (upper entries, then mirrored ones)
  std::vector<int32_t> mv_ptr(nk + 1, 0), mv;
  {
    std::vector<std::vector<int32_t>> up(nk), lo(nk);
    for (int32_t k = 0; k < nz; ++k) {
      up[ds->kkt_rows[k]].push_back(k);
      if (ds->kkt_rows[k] != ds->kkt_cols[k]) lo[ds->kkt_cols[k]].push_back(k);
    }
    for (int32_t o = 0; o < nk; ++o) {
      for (int32_t k : up[o]) mv.insert(mv.end(), {ds->kkt_cols[k], k});
      for (int32_t k : lo[o]) mv.insert(mv.end(), {ds->kkt_rows[k], k});
      mv_ptr[o + 1] = (int32_t)(mv.size() / 2);
    }
  }
  // rows of A (ascending column)
  std::vector<int32_t> arow_ptr(m + 1, 0), arow;
  {
    std::vector<std::vector<int32_t>> byrow(m);
    for (int32_t k = 0; k < ds->nnz_a; ++k) {
      WMPC_REQUIRE(ds->a_rows[k] >= 0 && ds->a_rows[k] < m && ds->a_cols[k] >= 0 && ds->a_cols[k] < n,
                   "a_rows/a_cols out of range");
      byrow[ds->a_rows[k]].push_back(k);
    }
    for (int32_t i = 0; i < m; ++i) {
      std::stable_sort(byrow[i].begin(), byrow[i].end(),
                       [&](int32_t a, int32_t b) { return ds->a_cols[a] < ds->a_cols[b]; });
      for (int32_t k : byrow[i]) arow.insert(arow.end(), {ds->a_cols[k], k});
      arow_ptr[i + 1] = (int32_t)(arow.size() / 2);
    }
  }
  // per constraint row: positions of its lower / upper multiplier and of its
  // equality multiplier, relative to the multiplier block of the M0 solution
  std::vector<int32_t> con(4 * (size_t)m, -1);
  for (int32_t i = 0; i < m; ++i) con[4 * i + 3] = 0;
  for (int32_t k = 0; k < nl; ++k) {
    const int32_t c = ds->lower_indices[k];
    WMPC_REQUIRE(c >= 0 && c < m && con[4 * c] < 0, "lower index out of range or repeated");
    con[4 * c] = k;
  }
  for (int32_t k = 0; k < nu; ++k) {
    const int32_t c = ds->upper_indices[k];
    WMPC_REQUIRE(c >= 0 && c < m && con[4 * c + 1] < 0, "upper index out of range or repeated");
    con[4 * c + 1] = nl + k;
  }
  for (int32_t e = 0; e < ne; ++e) {
    const int32_t c = ds->eq_indices[e];
    WMPC_REQUIRE(c >= 0 && c < m, "eq index out of range");
    WMPC_REQUIRE(con[4 * c] < 0 && con[4 * c + 1] < 0 && con[4 * c + 2] < 0,
                 "an equality row is also an inequality row");
    con[4 * c + 2] = nl + nu + e;
  }
  std::vector<int32_t> pent(2 * (size_t)ds->nnz_p), aent(2 * (size_t)ds->nnz_a);
  for (int32_t k = 0; k < ds->nnz_p; ++k) {
    pent[2 * k] = ds->p_upper_rows[k];
    pent[2 * k + 1] = ds->p_upper_cols[k];
  }
  for (int32_t k = 0; k < ds->nnz_a; ++k) {
    aent[2 * k] = ds->a_rows[k];
    aent[2 * k + 1] = ds->a_cols[k];
  }
  // the same rows in the sweeps' first-touch order, value sources inlined
  const std::vector<int32_t>& rperm = kkt->host.rperm;
  WMPC_REQUIRE((int32_t)rperm.size() == nk, "KKT plan has no first-touch order");
  std::vector<int32_t> mrow(2 * (size_t)nk), ment;
  ment.reserve(2 * mv.size());
  for (int32_t i = 0; i < nk; ++i) {
    const int32_t o = rperm[i];
    WMPC_REQUIRE(o >= 0 && o < nk, "first-touch order out of range");
    for (int32_t e = mv_ptr[o]; e < mv_ptr[o + 1]; ++e) {
      const int32_t k = mv[2 * (size_t)e + 1];
      ment.insert(ment.end(), {mv[2 * (size_t)e], slot[4 * (size_t)k], slot[4 * (size_t)k + 1],
                               slot[4 * (size_t)k + 3]});
    }
    mrow[2 * (size_t)i] = o;
    mrow[2 * (size_t)i + 1] = (int32_t)(ment.size() / 4);
  }
  Blob blob;
  const size_t t_slot = blob.add(slot), t_mvp = blob.add(mv_ptr), t_mv = blob.add(mv),
               t_arp = blob.add(arow_ptr), t_ar = blob.add(arow), t_con = blob.add(con),
               t_pe = blob.add(pent), t_ae = blob.add(aent), t_mr = blob.add(mrow),
               t_me = blob.add(ment);
  auto* plan = new wmpc_adjoint_plan();
  if (int rc = blob.upload(&plan->block, "adjoint KKT plan")) {
    delete plan;
    return rc;
  }
  plan->kkt_form = true;
  AdjKktDev& d = plan->k;
  d.n = n;
  d.m = m;
  d.nk = nk;
  d.nl = nl;
  d.ni = nl + nu;
  d.half = n + nl + nu + ne;
  d.nnz_p = ds->nnz_p;
  d.nnz_a = ds->nnz_a;
  d.delta = (float)ds->delta;
  d.slot = blob.at<int4>(t_slot);
  d.mv_ptr = blob.at<int32_t>(t_mvp);
  d.mv = blob.at<int2>(t_mv);
  d.arow_ptr = blob.at<int32_t>(t_arp);
  d.arow = blob.at<int2>(t_ar);
  d.con = blob.at<int4>(t_con);
  d.pent = blob.at<int2>(t_pe);
  d.aent = blob.at<int2>(t_ae);
  d.nmv = (int32_t)(ment.size() / 4);
  d.mrow = blob.at<int2>(t_mr);
  d.ment = blob.at<int4>(t_me);
  // the M0 width the reference's adjoint_derivative_compute returns
  plan->d.n = n;
  plan->d.m = m;
  plan->d.ni = nl + nu;
  plan->d.half = d.half;
  plan->d.dim = 2 * d.half;
  plan->adj = kkt;
  *out = plan;
  return WMPC_OK;
}

int32_t wmpc_adjoint_dim(const wmpc_adjoint_plan* plan) { return plan ? plan->d.dim : -1; }

int32_t wmpc_adjoint_is_kkt_form(const wmpc_adjoint_plan* plan) {
  return plan ? (plan->kkt_form ? 1 : 0) : -1;
}

int64_t wmpc_adjoint_work_floats(const wmpc_adjoint_plan* plan, int64_t ld) {
  if (!plan || ld < 0) return -1;
  if (plan->kkt_form)
    return ((int64_t)plan->adj->host.nnz_l + 5 * (int64_t)plan->k.nk + 3 * (int64_t)plan->k.m) * ld;
  return ((int64_t)plan->adj->host.nnz_l + 4 * (int64_t)plan->d.dim + 2 * (int64_t)plan->d.ni) * ld;
}

int wmpc_osqp_adjoint_f32(const wmpc_adjoint_plan* plan, const float* p_values,
                          const float* a_values, int64_t pm_ld, const float* l, const float* u,
                          const float* x, const float* y, const float* dx, const float* dy,
                          int32_t refine, float* work, int64_t work_len, float* sol, float* dq,
                          float* dl, float* du, float* dp, float* da, int64_t batch, int64_t ld,
                          void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && (batch == 0 || (p_values && a_values && l && u && x && y && dx && dy && work &&
                                       dq && dl && du)),
               "null argument");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  WMPC_REQUIRE(pm_ld == 0 || pm_ld >= batch, "pm_ld must be 0 (shared) or >= batch");
  WMPC_REQUIRE(refine >= 0, "refine must be nonnegative");
  if (batch == 0) return WMPC_OK;
  WMPC_REQUIRE(work_len >= wmpc_adjoint_work_floats(plan, ld),
               "work is smaller than wmpc_adjoint_work_floats(plan, ld)");
  const QdldlProg& P = plan->adj->prog;
  const size_t lds = adjoint_lds(P);
  if (lds > kMaxLdsBytes) {
    set_error("adjoint pattern needs more LDS than a workgroup can have");
    return WMPC_ERR_UNSUPPORTED;
  }
  const int32_t ipw = instances_per_wave(batch, P);
  if (plan->kkt_form) {
    const AdjKktDev& kv = plan->k;
    const int64_t nk = kv.nk;
    float* lx = work;
    float* dinv = lx + (int64_t)P.nnz_l * ld;
    float* stash = dinv + nk * ld;
    float* rbuf = stash + nk * ld;
    float* wv = rbuf + nk * ld;
    float* dg = wv + nk * ld;
    float* rc = dg + (int64_t)kv.m * ld;
    float* axv = rc + (int64_t)kv.m * ld;
    float* scv = axv + (int64_t)kv.m * ld;
    if (lds > 64 * 1024)
      WMPC_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_osqp_adjoint_kkt),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_osqp_adjoint_kkt, dim3((unsigned)wave_grid(batch, ipw)), dim3(kWave), lds,
                       (hipStream_t)stream, P, kv, p_values, a_values, pm_ld, l, u, x, y, dx, dy, lx,
                       dinv, stash, rbuf, wv, dg, rc, axv, scv, sol, dq, dl, du, dp, da, refine, batch,
                       ld, ipw);
    return check_launch("osqp adjoint (KKT form)");
  }
  const AdjDev& dv = plan->d;
  const int64_t dim = dv.dim, nl = P.nnz_l;
  float* lx = work;
  float* dinv = lx + nl * ld;
  float* stash = dinv + dim * ld;
  float* rbuf = stash + dim * ld;
  float* solw = rbuf + dim * ld;
  float* lam = solw + dim * ld;
  float* slk = lam + (int64_t)dv.ni * ld;
  if (lds > 64 * 1024)
    WMPC_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_osqp_adjoint),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_osqp_adjoint, dim3((unsigned)wave_grid(batch, ipw)), dim3(kWave), lds,
                     (hipStream_t)stream, P, dv, p_values, a_values, pm_ld, l, u, x, y, dx, dy, lx,
                     dinv, stash, rbuf, sol ? sol : solw, lam, slk, dq, dl, du, dp, da, refine, batch,
                     ld, ipw);
  return check_launch("osqp adjoint");
}

}  // extern "C"
