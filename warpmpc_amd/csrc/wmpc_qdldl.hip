// Batched fixed-pattern QDLDL factor / solve for MI355X (gfx950) behind the C ABI
// of include/warpmpc_hip.h.
//
// Replaces the Warp kernels of warpmpc/jax_qdldl/warp_backend.py:41-454 and the
// launch glue of :583-1076.  Launch shape: one lane per instance, 256-lane
// workgroups, batch-major data, reach slots of the current row in LDS.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wmpc_qdldl_device.h"

namespace wmpc {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// kernels

__global__ __launch_bounds__(kBlock) void k_qdldl_factor(QdldlTables t, DirectSource src,
                                                         float* __restrict__ lx,
                                                         float* __restrict__ d,
                                                         float* __restrict__ dinv, int64_t batch,
                                                         int64_t ld) {
  extern __shared__ float ys[];
  const int64_t b = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (b >= batch) return;
  factor_lane<DirectSource, false>(t, src, b, b, ld, lx, d, dinv, ys + threadIdx.x, kBlock,
                                   ArrayRhs{nullptr, 0}, nullptr);
}

__global__ __launch_bounds__(kBlock) void k_qdldl_factor_solve(
    QdldlTables t, DirectSource src, ArrayRhs rhs, float* __restrict__ x, float* __restrict__ lx,
    float* __restrict__ d, float* __restrict__ dinv, float* __restrict__ w, int64_t batch,
    int64_t ld) {
  extern __shared__ float ys[];
  const int64_t b = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (b >= batch) return;
  factor_lane<DirectSource, true>(t, src, b, b, ld, lx, d, dinv, ys + threadIdx.x, kBlock, rhs, w);
  backward_cols<false>(t, lx, dinv, ld, b, ld, w);
  for (int32_t i = 0; i < t.n; ++i) x[(int64_t)i * ld + b] = w[(int64_t)t.pinv[i] * ld + b];
}

template <bool SHARED>
__global__ __launch_bounds__(kBlock) void k_qdldl_solve(QdldlTables t,
                                                        const float* __restrict__ lx,
                                                        const float* __restrict__ dinv,
                                                        int64_t f_ld, ArrayRhs rhs,
                                                        float* __restrict__ x,
                                                        float* __restrict__ w, int64_t batch,
                                                        int64_t ld) {
  const int64_t b = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (b >= batch) return;
  forward_rows<SHARED>(t, lx, f_ld, b, ld, rhs, w);
  backward_cols<SHARED>(t, lx, dinv, f_ld, b, ld, w);
  for (int32_t i = 0; i < t.n; ++i) x[(int64_t)i * ld + b] = w[(int64_t)t.pinv[i] * ld + b];
}

// ---------------------------------------------------------------------------
// plan construction (host)

static int build_tables(const wmpc_qdldl_pattern* pt, std::vector<int32_t>& blob,
                        QdldlTables& t, std::vector<size_t>& offs) {
  const int32_t n = pt->n;
  WMPC_REQUIRE(n > 0 && pt->nnz_a > 0 && pt->nnz_l >= 0, "pattern sizes must be positive");
  WMPC_REQUIRE(pt->perm_to_orig && pt->diag_pos && pt->init_ptr && pt->init_rows &&
                   pt->init_pos && pt->proc_ptr && pt->proc_cols && pt->proc_lidx &&
                   pt->proc_off && pt->lp && pt->li && pt->p && pt->pinv,
               "pattern has a null table");
  WMPC_REQUIRE(pt->lp[0] == 0 && pt->lp[n] == pt->nnz_l, "lp must span [0, nnz_l]");
  const int32_t n_init = pt->init_ptr[n];
  const int32_t n_proc = pt->proc_ptr[n];
  WMPC_REQUIRE(n_proc == pt->nnz_l, "row programs must produce every entry of L once");
  std::vector<int32_t> diag_src(n), init_slot(n_init), init_src(n_init), upd_ptr(n_proc + 1);
  std::vector<int32_t> upd_slot;
  std::vector<int32_t> slot_of(n, -1), stamp(n, -1);
  int32_t max_reach = 1;
  for (int32_t k = 0; k < n; ++k) {
    const int32_t dp = pt->diag_pos[k];
    WMPC_REQUIRE(dp >= 0 && dp < pt->nnz_a, "diag_pos out of range");
    diag_src[k] = pt->perm_to_orig[dp];
    const int32_t r0 = pt->proc_ptr[k], r1 = pt->proc_ptr[k + 1];
    WMPC_REQUIRE(r0 <= r1, "proc_ptr must be nondecreasing");
    max_reach = std::max(max_reach, r1 - r0);
    for (int32_t r = r0; r < r1; ++r) {
      const int32_t c = pt->proc_cols[r];
      WMPC_REQUIRE(c >= 0 && c < k, "reach column must precede its row");
      slot_of[c] = r - r0;
      stamp[c] = k;
    }
    for (int32_t j = pt->init_ptr[k]; j < pt->init_ptr[k + 1]; ++j) {
      const int32_t row = pt->init_rows[j];
      WMPC_REQUIRE(row >= 0 && row < k && stamp[row] == k, "init row outside the reach of its row");
      init_slot[j] = slot_of[row];
      const int32_t ip = pt->init_pos[j];
      WMPC_REQUIRE(ip >= 0 && ip < pt->nnz_a, "init_pos out of range");
      init_src[j] = pt->perm_to_orig[ip];
    }
    for (int32_t r = r0; r < r1; ++r) {
      const int32_t c = pt->proc_cols[r];
      const int32_t off = pt->proc_off[r];
      WMPC_REQUIRE(off >= 0 && pt->lp[c] + off <= pt->lp[c + 1], "proc_off out of range");
      WMPC_REQUIRE(pt->proc_lidx[r] == pt->lp[c] + off, "proc_lidx must be lp[c] + offset");
      upd_ptr[r] = (int32_t)upd_slot.size();
      for (int32_t q = 0; q < off; ++q) {
        const int32_t row = pt->li[pt->lp[c] + q];
        WMPC_REQUIRE(row >= 0 && row < k && stamp[row] == k && slot_of[row] > r - r0,
                     "update target must be a later reach slot");
        upd_slot.push_back(slot_of[row]);
      }
    }
  }
  upd_ptr[n_proc] = (int32_t)upd_slot.size();
  for (int32_t i = 0; i < n; ++i) {
    WMPC_REQUIRE(pt->p[i] >= 0 && pt->p[i] < n && pt->pinv[pt->p[i]] == i, "p/pinv mismatch");
  }
  auto append = [&](const int32_t* src, size_t cnt) {
    offs.push_back(blob.size());
    blob.insert(blob.end(), src, src + cnt);
    while (blob.size() % 64) blob.push_back(0);  // 256-byte aligned sub-arrays
  };
  append(diag_src.data(), n);
  append(pt->init_ptr, n + 1);
  append(init_slot.data(), n_init);
  append(init_src.data(), n_init);
  append(pt->proc_ptr, n + 1);
  append(pt->proc_cols, n_proc);
  append(pt->proc_lidx, n_proc);
  append(upd_ptr.data(), n_proc + 1);
  append(upd_slot.data(), upd_slot.size());
  append(pt->lp, n + 1);
  append(pt->li, pt->nnz_l);
  append(pt->p, n);
  append(pt->pinv, n);
  t.n = n;
  t.nnz_a = pt->nnz_a;
  t.nnz_l = pt->nnz_l;
  t.max_reach = max_reach;
  t.n_proc = n_proc;
  t.n_upd = (int32_t)upd_slot.size();
  return WMPC_OK;
}

static void bind_tables(QdldlTables& t, const int32_t* base, const std::vector<size_t>& offs) {
  size_t i = 0;
  t.diag_src = base + offs[i++];
  t.init_ptr = base + offs[i++];
  t.init_slot = base + offs[i++];
  t.init_src = base + offs[i++];
  t.proc_ptr = base + offs[i++];
  t.proc_col = base + offs[i++];
  t.proc_lidx = base + offs[i++];
  t.upd_ptr = base + offs[i++];
  t.upd_slot = base + offs[i++];
  t.lp = base + offs[i++];
  t.li = base + offs[i++];
  t.p = base + offs[i++];
  t.pinv = base + offs[i++];
}

static int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + " launch: " + hipGetErrorString(e));
    return WMPC_ERR_HIP;
  }
  return WMPC_OK;
}

static size_t lds_bytes(const QdldlTables& t) { return (size_t)t.max_reach * kBlock * sizeof(float); }

}  // namespace wmpc

using namespace wmpc;

extern "C" {

const char* wmpc_last_error(void) { return g_last_error.c_str(); }
const char* wmpc_version(void) { return "warpmpc_hip 0.1.0 gfx950"; }

int wmpc_qdldl_plan_create(const wmpc_qdldl_pattern* pattern, wmpc_qdldl_plan** out) {
  clear_error();
  WMPC_REQUIRE(pattern && out, "null argument");
  *out = nullptr;
  std::vector<int32_t> blob;
  std::vector<size_t> offs;
  QdldlTables t;
  int rc = build_tables(pattern, blob, t, offs);
  if (rc != WMPC_OK) return rc;
  WMPC_REQUIRE(lds_bytes(t) <= 64 * 1024, "reach too wide for the LDS slot buffer");
  auto* plan = new wmpc_qdldl_plan();
  hipError_t e = hipGetDevice(&plan->device);
  if (e == hipSuccess) e = hipMalloc(&plan->block, blob.size() * sizeof(int32_t));
  if (e == hipSuccess)
    e = hipMemcpy(plan->block, blob.data(), blob.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("plan upload: ") + hipGetErrorString(e));
    if (plan->block) (void)hipFree(plan->block);
    delete plan;
    return WMPC_ERR_HIP;
  }
  plan->t = t;
  bind_tables(plan->t, static_cast<const int32_t*>(plan->block), offs);
  *out = plan;
  return WMPC_OK;
}

int wmpc_qdldl_plan_destroy(wmpc_qdldl_plan* plan) {
  clear_error();
  if (!plan) return WMPC_OK;
  if (plan->block) WMPC_HIP_CHECK(hipFree(plan->block));
  delete plan;
  return WMPC_OK;
}

int wmpc_qdldl_plan_info(const wmpc_qdldl_plan* plan, int32_t* n, int32_t* nnz_a, int32_t* nnz_l,
                         int32_t* max_reach) {
  clear_error();
  WMPC_REQUIRE(plan, "null plan");
  if (n) *n = plan->t.n;
  if (nnz_a) *nnz_a = plan->t.nnz_a;
  if (nnz_l) *nnz_l = plan->t.nnz_l;
  if (max_reach) *max_reach = plan->t.max_reach;
  return WMPC_OK;
}

int wmpc_qdldl_factor_f32(const wmpc_qdldl_plan* plan, const float* a_values, float* lx, float* d,
                          float* dinv, int64_t batch, int64_t ld, void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && (batch == 0 || (a_values && lx && d && dinv)), "null argument");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  if (batch == 0) return WMPC_OK;
  const int64_t grid = (batch + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_qdldl_factor, dim3((unsigned)grid), dim3(kBlock), lds_bytes(plan->t),
                     (hipStream_t)stream, plan->t, DirectSource{a_values, ld}, lx, d, dinv, batch,
                     ld);
  return check_launch("qdldl factor");
}

int wmpc_qdldl_solve_f32(const wmpc_qdldl_plan* plan, const float* lx, int64_t lx_ld,
                         const float* dinv, int64_t d_ld, const float* rhs, float* x, float* work,
                         int64_t batch, int64_t ld, void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && (batch == 0 || (lx && dinv && rhs && x && work)), "null argument");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  WMPC_REQUIRE((lx_ld == 0 && d_ld == 0) || (lx_ld >= batch && d_ld >= batch),
               "factor strides must both be 0 (shared) or both >= batch");
  WMPC_REQUIRE(lx_ld == d_ld, "lx and dinv must share one instance stride");
  if (batch == 0) return WMPC_OK;
  const int64_t grid = (batch + kBlock - 1) / kBlock;
  if (lx_ld == 0)
    hipLaunchKernelGGL(k_qdldl_solve<true>, dim3((unsigned)grid), dim3(kBlock), 0,
                       (hipStream_t)stream, plan->t, lx, dinv, (int64_t)0, ArrayRhs{rhs, ld}, x,
                       work, batch, ld);
  else
    hipLaunchKernelGGL(k_qdldl_solve<false>, dim3((unsigned)grid), dim3(kBlock), 0,
                       (hipStream_t)stream, plan->t, lx, dinv, lx_ld, ArrayRhs{rhs, ld}, x, work,
                       batch, ld);
  return check_launch("qdldl solve");
}

int wmpc_qdldl_factor_solve_f32(const wmpc_qdldl_plan* plan, const float* a_values,
                                const float* rhs, float* x, float* lx, float* d, float* dinv,
                                float* work, int64_t batch, int64_t ld, void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && (batch == 0 || (a_values && rhs && x && lx && d && dinv && work)), "null argument");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  if (batch == 0) return WMPC_OK;
  const int64_t grid = (batch + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_qdldl_factor_solve, dim3((unsigned)grid), dim3(kBlock),
                     lds_bytes(plan->t), (hipStream_t)stream, plan->t,
                     DirectSource{a_values, ld}, ArrayRhs{rhs, ld}, x, lx, d, dinv, work, batch,
                     ld);
  return check_launch("qdldl factor+solve");
}

}  // extern "C"
