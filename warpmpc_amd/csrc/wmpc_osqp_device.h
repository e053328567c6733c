// Device side of the fused OSQP solve (gfx950): the KKT value sources and
// update-pass tables of one pattern, the ADMM update pass and one instance's
// whole solve.  Shared by the single-pattern launch (wmpc_osqp.hip) and the
// heterogeneous multi-pattern launch (wmpc_hetero.hip).
#pragma once

#include <algorithm>

#include "wmpc_plan.h"
#include "wmpc_stream.h"

namespace wmpc {

struct OsqpDev {
  int32_t n = 0, m = 0, nnz_p = 0, nnz_a = 0, nnz_kkt = 0;
  float sigma = 0.f, alpha = 0.f;
  // KKT value sources, indexed by original KKT slot
  const int32_t* src_kind = nullptr;
  const int32_t* src_idx = nullptr;
  const float* src_scale = nullptr;
  const float* src_off = nullptr;
  // update-pass entries: variables {c, o, first-touch position, q scale},
  // constraints {c, i, first-touch position, rho⁻¹} + {rho, E, 0, 0}
  int32_t nvar = 0, ncon = 0;
  const int4* vtab = nullptr;
  const int4* ctab = nullptr;
  const float* sd = nullptr;      // D      [n]
  const float* sdinv = nullptr;   // D⁻¹    [n]
  const float* se = nullptr;      // E      [m]
  const float* seinv = nullptr;   // E⁻¹    [m]
  const float* y_scale = nullptr;    // [m]
  const float* y_unscale = nullptr;  // [m]
  // gather tables of the residuals: rows of A, columns of A, rows of full P
  const int32_t* arow_ptr = nullptr;  // [m+1]
  const int2* arow = nullptr;         // {col, A value slot}
  const int32_t* acol_ptr = nullptr;  // [n+1]
  const int2* acol = nullptr;         // {row, A value slot}
  const int32_t* prow_ptr = nullptr;  // [n+1]
  const int2* prow = nullptr;         // {col, P value slot}
};

// Tables of the one-wavefront factorization of a matrix shared by the batch
// (k_osqp_factor_shared): the up-looking recurrence of QDLDL with the updates
// of each reach entry spread over the lanes.  The kernel stages all of them in
// LDS; patterns whose tables do not fit (shared_factor_lds above the LDS of a
// workgroup) use the one-lane row-program kernel instead.
struct SharedFactorProg {
  int32_t n = 0, nnz_l = 0, nsteps = 0, nnz_kkt = 0, ninit = 0;
  // row k: {end of its reach steps, first init entry, end of init entries,
  // original KKT slot of K[k,k]}
  const int4* rows = nullptr;      // [n]
  // reach step: {c | (entries of column c before row k) << 16, first L slot of column c}
  const int2* steps = nullptr;     // [nsteps]
  const int32_t* li = nullptr;     // [nnz_l] L row indices
  const int2* init = nullptr;      // [ninit] {row, original KKT slot}: upper entries of column k
};

// LDS layout of k_osqp_factor_shared (float offsets): rows (4n), init entries
// (2 ninit), steps (2 nsteps), row workspace and D⁻¹ interleaved (2 (n + 1): a
// dummy row for idle lanes), L rows and values interleaved (2 (nnz_l + 1)).
// The KKT values live in the row and init records.  The 16- and 8-byte records
// come first, so they stay aligned.
__host__ __device__ inline int64_t shared_factor_lds_floats(const SharedFactorProg& F) {
  return 4 * (int64_t)F.n + 2 * (int64_t)F.ninit + 2 * (int64_t)F.nsteps + 2 * ((int64_t)F.n + 1) +
         2 * ((int64_t)F.nnz_l + 1);
}
inline size_t shared_factor_lds(const SharedFactorProg& F) {
  return (size_t)shared_factor_lds_floats(F) * sizeof(float);
}

}  // namespace wmpc

struct wmpc_osqp_plan {
  wmpc::OsqpDev d;
  wmpc::SharedFactorProg sf;
  const wmpc_qdldl_plan* kkt = nullptr;
  void* block = nullptr;
  // straight-line shared-factor ADMM kernel generated for this pattern
  // (hipModule_t / hipFunction_t of an attached code object, or null)
  void* spec_module = nullptr;
  void* spec_fn = nullptr;
};

namespace wmpc {

// Solution entry c of the backward sweep goes back into the stash.
struct StashOut {
  float* __restrict__ stash;
  int64_t ld, b;
  __device__ __forceinline__ void operator()(int32_t c, int32_t, float v) const {
    (stash + (int64_t)c * ld)[(uint32_t)b] = v;
  }
};

// One pass over the iterate (solver.py:806-820): unless FIRST, the relaxation,
// projection and dual update with the KKT solution S; unless LAST, the next
// right-hand side R (solver.py:808-810) at the entry's first-touch position.
// The entry tables are read as streams (TableStream): every lane takes part.
template <bool FIRST, bool LAST>
__device__ __forceinline__ void admm_pass(const OsqpDev& dv, const CStrided& S, const Strided& R,
                                          const Strided& X, const Strided& Z, const Strided& Y,
                                          const CStrided& Q, const CStrided& Lo,
                                          const CStrided& Up) {
  const float alpha = dv.alpha, oma = 1.0f - dv.alpha, sigma = dv.sigma;
  // entries per step: each entry reads and writes only its own variable or
  // constraint, so the operand loads of kU entries are issued before any of
  // their stores (several round trips in flight instead of one)
  constexpr int kU = 8;
  TableStream<int4> vt(dv.vtab, dv.nvar, 0);
  auto var = [&](const int4 t, float xo, float sx, float qv) {
    float xn = xo;
    if constexpr (!FIRST) {
      xn = alpha * sx + oma * xo;
      X(t.y) = xn;
    }
    if constexpr (!LAST) R(t.z) = sigma * xn - __int_as_float(t.w) * qv;
  };
  int32_t v = 0;
  for (; v + kU <= dv.nvar; v += kU) {
    int4 t[kU];
    float xo[kU], sx[kU], qv[kU];
#pragma unroll
    for (int i = 0; i < kU; ++i) t[i] = vt[v + i];
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      xo[i] = X(t[i].y);
      sx[i] = FIRST ? 0.f : S(t[i].x);
      qv[i] = LAST ? 0.f : Q(t[i].y);
    }
#pragma unroll
    for (int i = 0; i < kU; ++i) var(t[i], xo[i], sx[i], qv[i]);
  }
  for (; v < dv.nvar; ++v) {
    const int4 t = vt[v];
    var(t, X(t.y), FIRST ? 0.f : S(t.x), LAST ? 0.f : Q(t.y));
  }
  TableStream<int4> ct(dv.ctab, 2 * dv.ncon, 0);
  auto con = [&](const int4 t, const int4 t2, float zo, float yo, float sx, float lo, float up) {
    const float rinv = __int_as_float(t.w);
    float zn = zo, yn = yo;
    if constexpr (!FIRST) {
      const float rho = __int_as_float(t2.x), e = __int_as_float(t2.y);
      const float rz = zo - rinv * yo;
      const float zt = rz + rinv * sx;
      const float zr = alpha * zt + oma * zo;
      zn = nan_min(nan_max(zr + rinv * yo, lo * e), up * e);
      yn = yo + rho * (zr - zn);
      Z(t.y) = zn;
      Y(t.y) = yn;
    }
    if constexpr (!LAST) R(t.z) = zn - rinv * yn;
  };
  int32_t k = 0;
  for (; k + kU <= dv.ncon; k += kU) {
    int4 t[kU], t2[kU];
    float zo[kU], yo[kU], sx[kU], lo[kU], up[kU];
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      t[i] = ct[2 * (k + i)];
      t2[i] = FIRST ? t[i] : ct[2 * (k + i) + 1];
    }
#pragma unroll
    for (int i = 0; i < kU; ++i) {
      zo[i] = Z(t[i].y);
      yo[i] = Y(t[i].y);
      sx[i] = FIRST ? 0.f : S(t[i].x);
      lo[i] = FIRST ? 0.f : Lo(t[i].y);
      up[i] = FIRST ? 0.f : Up(t[i].y);
    }
#pragma unroll
    for (int i = 0; i < kU; ++i) con(t[i], t2[i], zo[i], yo[i], sx[i], lo[i], up[i]);
  }
  for (; k < dv.ncon; ++k) {
    const int4 t = ct[2 * k];
    const int4 t2 = FIRST ? t : ct[2 * k + 1];
    con(t, t2, Z(t.y), Y(t.y), FIRST ? 0.f : S(t.x), FIRST ? 0.f : Lo(t.y), FIRST ? 0.f : Up(t.y));
  }
}

// Scaled initial iterate of instance b (solver.py:842-847).
__device__ __forceinline__ void admm_initial(const OsqpDev& dv, const float* __restrict__ x0,
                                             const float* __restrict__ z0,
                                             const float* __restrict__ y0, const Strided& X,
                                             const Strided& Z, const Strided& Y, int64_t ld,
                                             int64_t b) {
  // V(i) = src[i] * f[i] (0 without src), 8 entries per step, loads first
  auto scaled = [&](const Strided& V, const float* __restrict__ src, const float* f, int32_t cnt) {
    if (!src) {
      for (int32_t i = 0; i < cnt; ++i) V(i) = 0.0f;
      return;
    }
    TableStream<float> fs(f, cnt, 0);
    int32_t i = 0;
    for (; i + 8 <= cnt; i += 8) {
      float v[8], g[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        v[j] = src[(int64_t)(i + j) * ld + b];
        g[j] = fs[i + j];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) V(i + j) = v[j] * g[j];
    }
    for (; i < cnt; ++i) V(i) = src[(int64_t)i * ld + b] * fs[i];
  };
  scaled(X, x0, dv.sdinv, dv.n);
  scaled(Z, z0, dv.se, dv.m);
  scaled(Y, y0, dv.y_scale, dv.m);
}

// Unscaling (solver.py:863-865), then the residuals and the objective with the
// unscaled P, A, q (solver.py:866-872), in gather form: no scatter scratch.
__device__ __forceinline__ void admm_finish(const OsqpDev& dv, const float* __restrict__ pv,
                                            const float* __restrict__ av, int64_t pm_ld,
                                            const CStrided& Q, const Strided& X, const Strided& Z,
                                            const Strided& Y, float* __restrict__ prim,
                                            float* __restrict__ dual, float* __restrict__ obj,
                                            int64_t b) {
  const int64_t pb = pm_ld ? b : 0;
  const int64_t pes = pm_ld ? pm_ld : 1;  // entry stride of P/A
  const int32_t n = dv.n, m = dv.m;
  // Tables are read as streams (every lane takes part) and each loop takes
  // several entries per step, their loads issued before their stores or
  // before the in-order sums that consume them.
  auto unscale = [&](const Strided& V, const float* f, int32_t cnt) {
    TableStream<float> fs(f, cnt, 0);
    int32_t i = 0;
    for (; i + 8 <= cnt; i += 8) {
      float v[8], g[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        v[j] = V(i + j);
        g[j] = fs[i + j];
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) V(i + j) = v[j] * g[j];
    }
    for (; i < cnt; ++i) V(i) *= fs[i];
  };
  unscale(X, dv.sd, n);
  unscale(Z, dv.seinv, m);
  unscale(Y, dv.y_unscale, m);
  // sum over one row's entries [e, end) of vals[slot] * V(col), in order
  auto row_sum = [&](TableStream<int2>& ents, int32_t& e, int32_t end, const float* vals,
                     const Strided& V) {
    float s = 0.0f;
    for (; e + 4 <= end; e += 4) {
      int2 t[4];
      float v[4], x[4];
      ents.get4(e, t[0], t[1], t[2], t[3]);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[j] = vals[(int64_t)t[j].y * pes + pb];
        x[j] = V(t[j].x);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) s += v[j] * x[j];
    }
    for (; e < end; ++e) {
      const int2 t = ents[e];
      s += vals[(int64_t)t.y * pes + pb] * V(t.x);
    }
    return s;
  };
  float pr = 0.0f;
  {
    TableStream<int32_t> ends(dv.arow_ptr + 1, m, 0);
    TableStream<int2> ents(dv.arow, dv.arow_ptr[m], 0);
    int32_t e = dv.arow_ptr[0];
    for (int32_t i = 0; i < m; ++i) {
      const float s = row_sum(ents, e, ends[i], av, X);
      pr = nan_max(pr, fabsf(s - Z(i)));
    }
  }
  float du = 0.0f, quad = 0.0f, lin = 0.0f;
  {
    TableStream<int32_t> pends(dv.prow_ptr + 1, n, 0), aends(dv.acol_ptr + 1, n, 0);
    TableStream<int2> pents(dv.prow, dv.prow_ptr[n], 0), aents(dv.acol, dv.acol_ptr[n], 0);
    int32_t ep = dv.prow_ptr[0], ea = dv.acol_ptr[0];
    for (int32_t j = 0; j < n; ++j) {
      const float px = row_sum(pents, ep, pends[j], pv, X);
      const float aty = row_sum(aents, ea, aends[j], av, Y);
      const float xj = X(j), qj = Q(j);
      du = nan_max(du, fabsf(px + qj + aty));
      quad += xj * px;
      lin += qj * xj;
    }
  }
  prim[b] = pr;
  dual[b] = du;
  obj[b] = 0.5f * quad + lin;
}

// One solve of instance b (solver.py:822-872) by this lane; every lane of the
// wavefront takes part (the sweeps stage operands wave-wide).
// SHARED: one factor for the batch ([nnz_l], [n_kkt]); FACTOR: assemble and
// factor each instance's KKT into the work area first.
template <bool SHARED, bool FACTOR>
__device__ __forceinline__ void admm_solve(
    const QdldlProg& P, const OsqpDev& dv, const float* __restrict__ pv,
    const float* __restrict__ av, int64_t pm_ld, const float* __restrict__ q,
    const float* __restrict__ l, const float* __restrict__ u, const float* __restrict__ x0,
    const float* __restrict__ z0, const float* __restrict__ y0, float* __restrict__ lx,
    float* __restrict__ dinv, int64_t f_ld, float* __restrict__ w, float* __restrict__ rbuf,
    float* __restrict__ x, float* __restrict__ z, float* __restrict__ y, float* __restrict__ prim,
    float* __restrict__ dual, float* __restrict__ obj, int32_t max_iter, int64_t ld, int64_t b,
    float* lds, int lane) {
  float* mywin = lds + (lane & (P.lcols - 1));
  const int64_t pb = pm_ld ? b : 0;
  const int64_t fb = SHARED ? 0 : b, fs = SHARED ? 1 : f_ld;
  if constexpr (FACTOR) {
    KktSource src{pv, av, pm_ld, dv.src_kind, dv.src_idx, dv.src_scale, dv.src_off};
    factor_rows<false>(P, src, pb, Strided{lx, fs, fb}, Strided{nullptr, 0, 0},
                       Strided{dinv, fs, fb}, mywin, NoRhs{}, Strided{nullptr, 0, 0});
  }
  const Strided X{x, ld, b}, Z{z, ld, b}, Y{y, ld, b}, R{rbuf, ld, b};
  const CStrided Q{q, ld, b}, Lo{l, ld, b}, Up{u, ld, b}, S{w, ld, b};
  admm_initial(dv, x0, z0, y0, X, Z, Y, ld, b);
  if (max_iter > 0) {
    const SweepProg SP = sweep_prog(P);
    admm_pass<true, false>(dv, S, R, X, Z, Y, Q, Lo, Up);
    for (int32_t it = 0; it < max_iter; ++it) {
      fwd_cols<SHARED>(SP, lx, dinv, fs, fb, rbuf, w, ld, b, lds, lane);
      bwd_cols<SHARED>(SP, lx, fs, fb, w, ld, b, lds, lane, StashOut{w, ld, b});
      if (it + 1 < max_iter)
        admm_pass<false, false>(dv, S, R, X, Z, Y, Q, Lo, Up);
      else
        admm_pass<false, true>(dv, S, R, X, Z, Y, Q, Lo, Up);
    }
  }
  admm_finish(dv, pv, av, pm_ld, Q, X, Z, Y, prim, dual, obj, b);
}

// LDS bytes of one ADMM workgroup (sweeps, plus the factor windows when the
// launch factors).
inline size_t admm_lds(const QdldlProg& P, bool shared, bool factor) {
  int64_t f = sweep_lds_floats(sweep_prog(P), shared);
  if (factor) f = std::max(f, factor_lds_floats(P, false));
  return (size_t)f * sizeof(float);
}

}  // namespace wmpc
