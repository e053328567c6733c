// QDLDL kernels that run directly on a plan's ragged tables (gfx950), behind
// include/warpmpc_hip.h:
//
// * float64 factor and solve (the Warp backend's float64 variant,
//   warpmpc/jax_qdldl/warp_backend.py:1211-1224): one lane per instance, the
//   up-looking recurrence of core.py:867-933 and the column sweeps of
//   core.py:935-1028 in that operation order, every working vector in global
//   memory, batch-major [entry][ld], so each access of a wavefront is one
//   contiguous stream across its instances;
// * the level-scheduled float32 solve (warp_backend.py:281-454, 970-1076):
//   several lanes per instance; the levels of core.py:488-510 in the
//   schedule of warp_backend.py:1110-1126 -- a level wider than the threshold
//   is one parallel step (its rows shared by the instance's lanes), a run of
//   narrower levels is a serial step (one lane) -- with the instance's
//   working vector in LDS.  The forward substitution gathers each row's
//   entries in ascending column order and the backward one each column's in
//   L order, the sequence of operations of the serial sweeps; with
//   floating-point contraction off in these kernels (no fused multiply-add)
//   the result equals the serial solve's bit for bit.
//
// The tables are uploaded with the plan (wmpc_qdldl_plan_create) into their
// own device block; they depend on the pattern only, not on the LDS budget.
#include <algorithm>
#include <string>
#include <vector>

#include "wmpc_blob.h"
#include "wmpc_plan.h"

namespace wmpc {

// Level tables: rows of L in row form (CSR: the column and L slot of each
// entry, ascending column), and the levels as index lists.
struct LevelProg {
  int32_t n = 0, nlev = 0;
  const int32_t* lev_ptr = nullptr;   // [nlev + 1]
  const int32_t* lev_rows = nullptr;  // [n] rows (= columns) by level
  const int32_t* row_ptr = nullptr;   // [n + 1]
  const int2* row_ent = nullptr;      // {column, L slot}
};

// Host side: levels of core.py:488-510 on the plan's L pattern.
static void build_levels(const wmpc_qdldl_pattern& pt, std::vector<int32_t>& lev_ptr,
                         std::vector<int32_t>& lev_rows, std::vector<int32_t>& row_ptr,
                         std::vector<int32_t>& row_ent) {
  const int32_t n = pt.n;
  std::vector<int32_t> level(n, 0), count(n + 1, 0);
  for (int32_t c = 0; c < n; ++c)
    for (int32_t q = pt.lp[c]; q < pt.lp[c + 1]; ++q) {
      const int32_t r = pt.li[q];
      level[r] = std::max(level[r], level[c] + 1);
      ++count[r + 1];
    }
  const int32_t nlev = n ? *std::max_element(level.begin(), level.end()) + 1 : 0;
  lev_ptr.assign(nlev + 1, 0);
  for (int32_t i = 0; i < n; ++i) ++lev_ptr[level[i] + 1];
  for (int32_t l = 0; l < nlev; ++l) lev_ptr[l + 1] += lev_ptr[l];
  lev_rows.assign(n, 0);
  std::vector<int32_t> fill(lev_ptr.begin(), lev_ptr.end() - (nlev ? 1 : 0));
  for (int32_t i = 0; i < n; ++i) lev_rows[fill[level[i]]++] = i;  // ascending within a level
  row_ptr.assign(n + 1, 0);
  for (int32_t i = 0; i < n; ++i) row_ptr[i + 1] = row_ptr[i] + count[i + 1];
  row_ent.assign(2 * (size_t)row_ptr[n], 0);
  std::vector<int32_t> at(row_ptr.begin(), row_ptr.end() - 1);
  for (int32_t c = 0; c < n; ++c)  // ascending column within each row
    for (int32_t q = pt.lp[c]; q < pt.lp[c + 1]; ++q) {
      const int32_t r = pt.li[q];
      row_ent[2 * (size_t)at[r]] = c;
      row_ent[2 * (size_t)at[r] + 1] = q;
      ++at[r];
    }
}

struct RawQdldl {
  int32_t n = 0, nnz_a = 0, nnz_l = 0;
  const int32_t* a_diag = nullptr;   // [n] original slot of A_kk (permuted k)
  const int32_t* init_ptr = nullptr; // [n + 1]
  const int2* init = nullptr;        // {row, original slot}
  const int32_t* proc_ptr = nullptr; // [n + 1]
  const int4* proc = nullptr;        // {column, L slot of L[k, c], entries of c before row k, 0}
  const int32_t* lp = nullptr;       // [n + 1]
  const int32_t* li = nullptr;       // [nnz_l]
  const int32_t* p = nullptr;        // [n]
  const int32_t* pinv = nullptr;     // [n]
  LevelProg lev;
};

}  // namespace wmpc

// device copies of one plan's raw tables (created with the plan)
struct wmpc_qdldl_raw {
  wmpc::RawQdldl r;
  void* block = nullptr;
};

namespace wmpc {

int raw_upload(const wmpc_qdldl_pattern& pt, wmpc_qdldl_raw** out) {
  const int32_t n = pt.n;
  std::vector<int32_t> a_diag(n), init, proc;
  for (int32_t k = 0; k < n; ++k) a_diag[k] = pt.perm_to_orig[pt.diag_pos[k]];
  for (int32_t j = 0; j < pt.init_ptr[n]; ++j) init.insert(init.end(), {pt.init_rows[j], pt.perm_to_orig[pt.init_pos[j]]});
  for (int32_t j = 0; j < pt.proc_ptr[n]; ++j)
    proc.insert(proc.end(), {pt.proc_cols[j], pt.proc_lidx[j], pt.proc_off[j], 0});
  std::vector<int32_t> lev_ptr, lev_rows, row_ptr, row_ent;
  build_levels(pt, lev_ptr, lev_rows, row_ptr, row_ent);
  Blob blob;
  const size_t t_ad = blob.add(a_diag), t_ip = blob.add(pt.init_ptr, (size_t)n + 1), t_in = blob.add(init),
               t_pp = blob.add(pt.proc_ptr, (size_t)n + 1), t_pr = blob.add(proc),
               t_lp = blob.add(pt.lp, (size_t)n + 1), t_li = blob.add(pt.li, (size_t)pt.nnz_l),
               t_p = blob.add(pt.p, (size_t)n), t_pi = blob.add(pt.pinv, (size_t)n), t_lv = blob.add(lev_ptr),
               t_lr = blob.add(lev_rows), t_rp = blob.add(row_ptr), t_re = blob.add(row_ent);
  auto* raw = new wmpc_qdldl_raw();
  if (int rc = blob.upload(&raw->block, "raw QDLDL tables")) {
    delete raw;
    return rc;
  }
  RawQdldl& r = raw->r;
  r.n = n;
  r.nnz_a = pt.nnz_a;
  r.nnz_l = pt.nnz_l;
  r.a_diag = blob.at<int32_t>(t_ad);
  r.init_ptr = blob.at<int32_t>(t_ip);
  r.init = blob.at<int2>(t_in);
  r.proc_ptr = blob.at<int32_t>(t_pp);
  r.proc = blob.at<int4>(t_pr);
  r.lp = blob.at<int32_t>(t_lp);
  r.li = blob.at<int32_t>(t_li);
  r.p = blob.at<int32_t>(t_p);
  r.pinv = blob.at<int32_t>(t_pi);
  r.lev.n = n;
  r.lev.nlev = (int32_t)lev_ptr.size() - 1;
  r.lev.lev_ptr = blob.at<int32_t>(t_lv);
  r.lev.lev_rows = blob.at<int32_t>(t_lr);
  r.lev.row_ptr = blob.at<int32_t>(t_rp);
  r.lev.row_ent = blob.at<int2>(t_re);
  *out = raw;
  return WMPC_OK;
}

void raw_destroy(wmpc_qdldl_raw* raw) {
  if (!raw) return;
  if (raw->block) (void)hipFree(raw->block);
  delete raw;
}

int32_t raw_levels(const wmpc_qdldl_raw* raw) { return raw ? raw->r.lev.nlev : -1; }

template <class T>
struct G {  // batch-major global vector of one instance
  T* p;
  int64_t ld;
  int64_t b;
  __device__ __forceinline__ T& operator()(int64_t i) const { return p[i * ld + b]; }
};

// float64 factorization (core.py:867-933 order); y: [n][ld] scratch.
__global__ __launch_bounds__(64) void k_qdldl_factor_f64(RawQdldl R, const double* __restrict__ a,
                                                         double* __restrict__ lx, double* __restrict__ d,
                                                         double* __restrict__ dinv, double* __restrict__ yw,
                                                         int64_t batch, int64_t ld) {
#pragma clang fp contract(off)
  const int64_t b = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (b >= batch) return;
  const G<double> Y{yw, ld, b}, LX{lx, ld, b}, DI{dinv, ld, b};
  const G<const double> A{a, ld, b};
  for (int32_t i = 0; i < R.n; ++i) Y(i) = 0.0;
  for (int32_t k = 0; k < R.n; ++k) {
    double dk = A(R.a_diag[k]);
    for (int32_t j = R.init_ptr[k]; j < R.init_ptr[k + 1]; ++j) {
      const int2 e = R.init[j];
      Y(e.x) = Y(e.x) + A(e.y);
    }
    for (int32_t j = R.proc_ptr[k]; j < R.proc_ptr[k + 1]; ++j) {
      const int4 e = R.proc[j];
      const double val = Y(e.x);
      const int32_t base = R.lp[e.x];
      for (int32_t q = 0; q < e.z; ++q) {
        const double upd = -(LX(base + q) * val);
        Y(R.li[base + q]) = Y(R.li[base + q]) + upd;
      }
      const double lval = val * DI(e.x);
      LX(e.y) = lval;
      dk = dk - val * lval;
    }
    if (d) G<double>{d, ld, b}(k) = dk;
    DI(k) = 1.0 / dk;
    for (int32_t j = R.proc_ptr[k]; j < R.proc_ptr[k + 1]; ++j) Y(R.proc[j].x) = 0.0;
  }
}

// float64 solve x = Pᵀ L⁻ᵀ D⁻¹ L⁻¹ P rhs (core.py:935-1028 order); w: [n][ld]
// scratch; f_ld = 0 shares one factor.
__global__ __launch_bounds__(64) void k_qdldl_solve_f64(RawQdldl R, const double* __restrict__ lx,
                                                        const double* __restrict__ dinv, int64_t f_ld,
                                                        const double* __restrict__ rhs, double* __restrict__ x,
                                                        double* __restrict__ ww, int64_t batch, int64_t ld) {
#pragma clang fp contract(off)
  const int64_t b = (int64_t)blockIdx.x * 64 + threadIdx.x;
  if (b >= batch) return;
  const G<double> W{ww, ld, b};
  const G<const double> LX{lx, f_ld ? f_ld : 1, f_ld ? b : 0}, DI{dinv, f_ld ? f_ld : 1, f_ld ? b : 0};
  for (int32_t i = 0; i < R.n; ++i) W(i) = rhs[(int64_t)R.p[i] * ld + b];
  for (int32_t c = 0; c < R.n; ++c) {
    const double val = W(c);
    for (int32_t q = R.lp[c]; q < R.lp[c + 1]; ++q) {
      const double upd = -(LX(q) * val);
      W(R.li[q]) = W(R.li[q]) + upd;
    }
  }
  for (int32_t i = 0; i < R.n; ++i) W(i) = W(i) * DI(i);
  for (int32_t c = R.n - 1; c >= 0; --c) {
    double corr = 0.0;
    for (int32_t q = R.lp[c]; q < R.lp[c + 1]; ++q) corr = corr + LX(q) * W(R.li[q]);
    W(c) = W(c) - corr;
  }
  for (int32_t i = 0; i < R.n; ++i) x[(int64_t)i * ld + b] = W(R.pinv[i]);
}

// Level-scheduled float32 solve.  A wavefront serves 64 / lpi instances, lane
// l working on instance slot l / lpi as sub-lane l % lpi; the instance's
// vector lives in LDS [n][64 / lpi].  Every lane reaches every barrier.
__global__ __launch_bounds__(64) void k_qdldl_solve_levels(RawQdldl R, const float* __restrict__ lx,
                                                           const float* __restrict__ dinv, int64_t f_ld,
                                                           const float* __restrict__ rhs, float* __restrict__ x,
                                                           int32_t threshold, int32_t lpi, int64_t batch,
                                                           int64_t ld) {
  // no fused multiply-adds: the serial sweeps' rounding, operation by operation
#pragma clang fp contract(off)
  extern __shared__ float xs[];
  const int32_t g = 64 / lpi;
  const int32_t slot = threadIdx.x / lpi, sub = threadIdx.x % lpi;
  const int64_t b0 = (int64_t)blockIdx.x * g + slot;
  const int64_t b = b0 < batch ? b0 : batch - 1;
  const int64_t fs = f_ld ? f_ld : 1;
  const int64_t fb = f_ld ? b : 0;
  float* X = xs + slot;  // X[i * g]
  const int32_t n = R.n;
  for (int32_t i = sub; i < n; i += lpi) X[i * g] = rhs[(int64_t)R.p[i] * ld + b];
  __syncthreads();
  const LevelProg& V = R.lev;
  // forward substitution by levels (rows gather their columns)
  for (int32_t l = 0; l < V.nlev;) {
    const int32_t w = V.lev_ptr[l + 1] - V.lev_ptr[l];
    int32_t e = l + 1, first = sub, step = lpi;
    if (w <= threshold) {  // serial run of narrow levels: one lane
      while (e < V.nlev && V.lev_ptr[e + 1] - V.lev_ptr[e] <= threshold) ++e;
      first = sub == 0 ? 0 : V.lev_ptr[e] - V.lev_ptr[l];
      step = 1;
    }
    for (int32_t j = V.lev_ptr[l] + first; j < V.lev_ptr[e]; j += step) {
      const int32_t i = V.lev_rows[j];
      float s = X[i * g];
      for (int32_t k = V.row_ptr[i]; k < V.row_ptr[i + 1]; ++k) {
        const int2 t = V.row_ent[k];
        s = s + -(lx[(int64_t)t.y * fs + fb] * X[t.x * g]);
      }
      X[i * g] = s;
    }
    __syncthreads();
    l = e;
  }
  for (int32_t i = sub; i < n; i += lpi) X[i * g] = X[i * g] * dinv[(int64_t)i * fs + fb];
  __syncthreads();
  // backward substitution by levels in reverse (columns gather their rows)
  for (int32_t l = V.nlev - 1; l >= 0;) {
    const int32_t w = V.lev_ptr[l + 1] - V.lev_ptr[l];
    int32_t s0 = l, first = sub, step = lpi;
    if (w <= threshold) {
      while (s0 > 0 && V.lev_ptr[s0] - V.lev_ptr[s0 - 1] <= threshold) --s0;
      first = sub == 0 ? 0 : V.lev_ptr[l + 1] - V.lev_ptr[s0];
      step = 1;
    }
    // columns of levels s0..l, from the last level back (a column's rows lie
    // in later levels); within a serial run one lane walks them in reverse
    const int32_t lo = V.lev_ptr[s0], hi = V.lev_ptr[l + 1];
    for (int32_t j = hi - 1 - first; j >= lo; j -= step) {
      const int32_t c = V.lev_rows[j];
      float corr = 0.0f;
      for (int32_t q = R.lp[c]; q < R.lp[c + 1]; ++q)
        corr = corr + lx[(int64_t)q * fs + fb] * X[R.li[q] * g];
      X[c * g] = X[c * g] - corr;
    }
    __syncthreads();
    l = s0 - 1;
  }
  if (b0 < batch)
    for (int32_t i = sub; i < n; i += lpi) x[(int64_t)i * ld + b] = X[R.pinv[i] * g];
}

}  // namespace wmpc

using namespace wmpc;

static int raw_of(const wmpc_qdldl_plan* plan, const RawQdldl** r) {
  WMPC_REQUIRE(plan, "null plan");
  WMPC_REQUIRE(plan->raw, "host-only plan: create it with wmpc_qdldl_plan_create");
  *r = &plan->raw->r;
  return WMPC_OK;
}

extern "C" {

int64_t wmpc_qdldl_f64_work_doubles(const wmpc_qdldl_plan* plan, int64_t ld) {
  if (!plan || ld < 0) return -1;
  return (int64_t)plan->pattern.n * ld;
}

int wmpc_qdldl_factor_f64(const wmpc_qdldl_plan* plan, const double* a_values, double* lx, double* d,
                          double* dinv, double* work, int64_t work_len, int64_t batch, int64_t ld,
                          void* stream) {
  clear_error();
  const RawQdldl* R = nullptr;
  if (int rc = raw_of(plan, &R)) return rc;
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  if (batch == 0) return WMPC_OK;
  WMPC_REQUIRE(a_values && lx && dinv && work, "null argument");
  WMPC_REQUIRE(work_len >= wmpc_qdldl_f64_work_doubles(plan, ld), "work is smaller than n * ld doubles");
  hipLaunchKernelGGL(k_qdldl_factor_f64, dim3((unsigned)((batch + 63) / 64)), dim3(64), 0, (hipStream_t)stream,
                     *R, a_values, lx, d, dinv, work, batch, ld);
  return check_launch("qdldl factor f64");
}

int wmpc_qdldl_solve_f64(const wmpc_qdldl_plan* plan, const double* lx, int64_t lx_ld, const double* dinv,
                         int64_t d_ld, const double* rhs, double* x, double* work, int64_t work_len,
                         int64_t batch, int64_t ld, void* stream) {
  clear_error();
  const RawQdldl* R = nullptr;
  if (int rc = raw_of(plan, &R)) return rc;
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  if (batch == 0) return WMPC_OK;
  WMPC_REQUIRE(lx && dinv && rhs && x && work, "null argument");
  WMPC_REQUIRE(lx_ld == d_ld && (lx_ld == 0 || lx_ld >= batch),
               "factor strides must be equal, 0 (shared) or >= batch");
  WMPC_REQUIRE(work_len >= wmpc_qdldl_f64_work_doubles(plan, ld), "work is smaller than n * ld doubles");
  hipLaunchKernelGGL(k_qdldl_solve_f64, dim3((unsigned)((batch + 63) / 64)), dim3(64), 0, (hipStream_t)stream,
                     *R, lx, dinv, lx_ld, rhs, x, work, batch, ld);
  return check_launch("qdldl solve f64");
}

int32_t wmpc_qdldl_plan_levels(const wmpc_qdldl_plan* plan) {
  return plan && plan->raw ? raw_levels(plan->raw) : -1;
}

int wmpc_qdldl_solve_levels_f32(const wmpc_qdldl_plan* plan, int32_t threshold, int32_t lanes_per_instance,
                                const float* lx, int64_t lx_ld, const float* dinv, int64_t d_ld,
                                const float* rhs, float* x, int64_t batch, int64_t ld, void* stream) {
  clear_error();
  const RawQdldl* R = nullptr;
  if (int rc = raw_of(plan, &R)) return rc;
  WMPC_REQUIRE(threshold >= 1, "threshold must be at least 1");
  const int32_t lpi = lanes_per_instance;
  WMPC_REQUIRE(lpi == 1 || lpi == 2 || lpi == 4 || lpi == 8 || lpi == 16 || lpi == 32 || lpi == 64,
               "lanes per instance must be a power of two in [1, 64]");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  if (batch == 0) return WMPC_OK;
  WMPC_REQUIRE(lx && dinv && rhs && x, "null argument");
  WMPC_REQUIRE(x != rhs, "x must not alias rhs");
  WMPC_REQUIRE(lx_ld == d_ld && (lx_ld == 0 || lx_ld >= batch),
               "factor strides must be equal, 0 (shared) or >= batch");
  const int32_t g = 64 / lpi;
  const size_t lds = (size_t)R->n * g * sizeof(float);
  WMPC_REQUIRE(lds <= kMaxLdsBytes, "n * (64 / lanes_per_instance) floats exceed a workgroup's LDS: use more lanes");
  if (lds > 64 * 1024)
    WMPC_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_qdldl_solve_levels),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_qdldl_solve_levels, dim3((unsigned)((batch + g - 1) / g)), dim3(64), lds,
                     (hipStream_t)stream, *R, lx, dinv, lx_ld, rhs, x, threshold, lpi, batch, ld);
  return check_launch("qdldl level-scheduled solve");
}

}  // extern "C"
