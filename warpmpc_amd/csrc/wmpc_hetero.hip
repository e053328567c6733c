// Heterogeneous batch of fixed-pattern OSQP problems in one persistent launch
// (gfx950), behind include/warpmpc_hip.h.
//
// Up to kMaxHeteroPatterns patterns, each with its own plan and batch.  The
// work of pattern p is cut into items of ipw_p instances (one wavefront each,
// ipw_p = the pattern's LDS columns, at most 64); items are numbered pattern
// by pattern in decreasing order of the pattern's L size, so the costliest
// items start first.  Each wavefront of the persistent grid takes the next
// item from a counter in global memory (one atomic add per item, by lane 0)
// until all are taken, and runs the same per-instance solve as the
// single-pattern launch (admm_solve: KKT assembly and factorization,
// max_iter ADMM iterations, unscaling, residuals).  Every wavefront reaches
// the loop's exit once the counter passes the item count.
//
// The patterns' device tables live in one plan-owned array; the per-call input
// and output pointers travel in the kernel arguments.
#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "wmpc_osqp_device.h"

namespace wmpc {

constexpr int kMaxHeteroPatterns = 12;

struct HeteroPattern {
  QdldlProg P;
  OsqpDev dv;
};

// One pattern's arrays for one call, in launch order.
struct HeteroIO {
  const float* p;
  const float* a;
  const float* q;
  const float* l;
  const float* u;
  const float* x0;
  const float* z0;
  const float* y0;
  float* x;
  float* z;
  float* y;
  float* prim;
  float* dual;
  float* obj;
  float* stash;
  float* rbuf;
  float* lx;
  float* dinv;
  int64_t pm_ld, batch, ld;
  int32_t pattern;    // index into the plan's pattern array
  int32_t ipw;        // instances per item
  int32_t items_end;  // items of this and every earlier entry
  int32_t pad;
};

struct HeteroArgs {
  HeteroIO io[kMaxHeteroPatterns];
  int32_t count;  // entries in io
  int32_t total;  // items over all entries
};

}  // namespace wmpc

struct wmpc_osqp_hetero_plan {
  std::vector<const wmpc_osqp_plan*> plans;
  wmpc::HeteroPattern* dev = nullptr;  // device copy of the patterns' tables
  size_t lds = 0;                      // LDS bytes of the launch (largest pattern)
  std::vector<int32_t> order;          // launch order: decreasing nnz_l
  std::vector<uint64_t> generation;    // of each pattern's KKT programs when copied
};

namespace wmpc {

__global__ __launch_bounds__(kWave) void k_counter_zero(int32_t* counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

__global__ __launch_bounds__(kWave) void k_osqp_hetero(const HeteroPattern* __restrict__ pats,
                                                       HeteroArgs A, int32_t* __restrict__ counter,
                                                       int32_t max_iter) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  for (;;) {
    int32_t item = 0;
    if (lane == 0) item = atomicAdd(counter, 1);
    item = __builtin_amdgcn_readfirstlane(item);
    if (item >= A.total) break;
    int32_t e = 0;
    while (A.io[e].items_end <= item) ++e;
    const HeteroIO& io = A.io[e];
    const int32_t begin = e ? A.io[e - 1].items_end : 0;
    const int64_t i = (int64_t)(item - begin) * io.ipw + (lane & (io.ipw - 1));
    const int64_t b = i < io.batch ? i : io.batch - 1;
    const HeteroPattern& hp = pats[io.pattern];
    admm_solve<false, true>(hp.P, hp.dv, io.p, io.a, io.pm_ld, io.q, io.l, io.u, io.x0, io.z0,
                            io.y0, io.lx, io.dinv, io.ld, io.stash, io.rbuf, io.x, io.z, io.y,
                            io.prim, io.dual, io.obj, max_iter, io.ld, b, lds, lane);
  }
}

// out[i][b] = in[rows[i]][b]; a source row outside [0, in_rows) gives NaN.
__global__ __launch_bounds__(256) void k_gather_rows(const int32_t* __restrict__ rows,
                                                     int32_t in_rows, const float* __restrict__ in,
                                                     int64_t in_ld, float* __restrict__ out,
                                                     int64_t out_ld, int64_t batch) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= batch) return;
  const int32_t i = blockIdx.y;
  const int32_t r = rows[i];
  out[(int64_t)i * out_ld + b] = (r >= 0 && r < in_rows) ? in[(int64_t)r * in_ld + b] : __int_as_float(0x7fc00000);
}

// Scratch floats of one pattern at instance stride ld: stash, staged
// right-hand side, per-instance L and D⁻¹; each region 64-float aligned.
static int64_t hetero_pattern_floats(const wmpc_osqp_plan* p, int64_t ld) {
  const int64_t nk = p->kkt->host.n, nl = p->kkt->host.nnz_l;
  auto al = [](int64_t v) { return (v + 63) / 64 * 64; };
  return 3 * al(nk * ld) + al(nl * ld);
}

}  // namespace wmpc

using namespace wmpc;

extern "C" {

int wmpc_osqp_hetero_plan_create(const wmpc_osqp_plan* const* plans, int32_t count,
                                 wmpc_osqp_hetero_plan** out) {
  clear_error();
  WMPC_REQUIRE(plans && out, "null argument");
  *out = nullptr;
  WMPC_REQUIRE(count >= 1 && count <= kMaxHeteroPatterns, "need 1 to 12 patterns");
  std::vector<HeteroPattern> host(count);
  size_t lds = 0;
  for (int32_t k = 0; k < count; ++k) {
    WMPC_REQUIRE(plans[k] && plans[k]->kkt && plans[k]->kkt->block, "null or host-only plan");
    host[k].P = plans[k]->kkt->prog;
    host[k].dv = plans[k]->d;
    lds = std::max(lds, admm_lds(host[k].P, false, true));
  }
  WMPC_REQUIRE(lds <= kMaxLdsBytes, "a pattern needs more LDS than a workgroup can have");
  auto* hp = new wmpc_osqp_hetero_plan();
  hp->plans.assign(plans, plans + count);
  for (int32_t k = 0; k < count; ++k) hp->generation.push_back(plans[k]->kkt->generation);
  hp->lds = lds;
  hp->order.resize(count);
  std::iota(hp->order.begin(), hp->order.end(), 0);
  std::stable_sort(hp->order.begin(), hp->order.end(), [&](int32_t a, int32_t b) {
    return host[a].P.nnz_l > host[b].P.nnz_l;
  });
  void* dev = nullptr;
  hipError_t e = hipMalloc(&dev, sizeof(HeteroPattern) * count);
  if (e == hipSuccess)
    e = hipMemcpy(dev, host.data(), sizeof(HeteroPattern) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("hetero plan upload: ") + hipGetErrorString(e));
    if (dev) (void)hipFree(dev);
    delete hp;
    return WMPC_ERR_HIP;
  }
  hp->dev = static_cast<HeteroPattern*>(dev);
  *out = hp;
  return WMPC_OK;
}

int wmpc_osqp_hetero_plan_destroy(wmpc_osqp_hetero_plan* plan) {
  clear_error();
  if (!plan) return WMPC_OK;
  if (plan->dev) WMPC_HIP_CHECK(hipFree(plan->dev));
  delete plan;
  return WMPC_OK;
}

int64_t wmpc_osqp_hetero_work_floats(const wmpc_osqp_hetero_plan* plan, const int64_t* ld) {
  if (!plan || !ld) return -1;
  int64_t total = 64;  // the item counter
  for (size_t k = 0; k < plan->plans.size(); ++k) {
    if (ld[k] < 0) return -1;
    total += hetero_pattern_floats(plan->plans[k], ld[k]);
  }
  return total;
}

int wmpc_osqp_solve_hetero_f32(const wmpc_osqp_hetero_plan* plan, const wmpc_osqp_hetero_io* io,
                               int32_t max_iter, float* work, int64_t work_len, void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && io && work, "null argument");
  WMPC_REQUIRE(max_iter >= 0, "max_iter must be nonnegative");
  const int32_t count = (int32_t)plan->plans.size();
  for (int32_t k = 0; k < count; ++k)
    WMPC_REQUIRE(plan->plans[k]->kkt->generation == plan->generation[k],
                 "a pattern's programs were re-derived (LDS budget) after the heterogeneous plan was "
                 "created: create it again");
  std::vector<int64_t> lds(count);
  for (int32_t k = 0; k < count; ++k) lds[k] = io[k].ld;
  WMPC_REQUIRE(work_len >= wmpc_osqp_hetero_work_floats(plan, lds.data()),
               "work is smaller than wmpc_osqp_hetero_work_floats(plan, ld)");
  // scratch regions in pattern order
  std::vector<float*> region(count);
  float* cursor = work + 64;
  for (int32_t k = 0; k < count; ++k) {
    region[k] = cursor;
    cursor += hetero_pattern_floats(plan->plans[k], io[k].ld);
  }
  HeteroArgs A{};
  int32_t items = 0;
  for (int32_t j = 0; j < count; ++j) {
    const int32_t k = plan->order[j];
    const wmpc_osqp_hetero_io& s = io[k];
    WMPC_REQUIRE(s.batch >= 0 && s.ld >= s.batch, "need 0 <= batch <= ld for every pattern");
    WMPC_REQUIRE(s.pm_ld == 0 || s.pm_ld >= s.batch, "pm_ld must be 0 (shared) or >= batch");
    WMPC_REQUIRE(s.batch == 0 || (s.p_values && s.a_values && s.q && s.l && s.u && s.x && s.z &&
                                  s.y && s.prim && s.dual && s.obj),
                 "null argument in a pattern's arrays");
    WMPC_REQUIRE((s.x0 == nullptr) == (s.z0 == nullptr) && (s.z0 == nullptr) == (s.y0 == nullptr),
                 "a warm start needs x0, z0 and y0 together");
    const QdldlProg& P = plan->plans[k]->kkt->prog;
    const int32_t ipw = std::min<int32_t>(kWave, P.lcols);
    const int64_t nitems = (s.batch + ipw - 1) / ipw;
    WMPC_REQUIRE((int64_t)items + nitems < INT32_MAX, "too many instances for one launch");
    items += (int32_t)nitems;
    const int64_t nk = P.n, nl = P.nnz_l;
    auto al = [](int64_t v) { return (v + 63) / 64 * 64; };
    HeteroIO& h = A.io[j];
    h.p = s.p_values;
    h.a = s.a_values;
    h.q = s.q;
    h.l = s.l;
    h.u = s.u;
    h.x0 = s.x0;
    h.z0 = s.z0;
    h.y0 = s.y0;
    h.x = s.x;
    h.z = s.z;
    h.y = s.y;
    h.prim = s.prim;
    h.dual = s.dual;
    h.obj = s.obj;
    h.stash = region[k];
    h.rbuf = h.stash + al(nk * s.ld);
    h.dinv = h.rbuf + al(nk * s.ld);
    h.lx = h.dinv + al(nk * s.ld);
    (void)nl;
    h.pm_ld = s.pm_ld;
    h.batch = s.batch;
    h.ld = s.ld;
    h.pattern = k;
    h.ipw = ipw;
    h.items_end = items;
  }
  A.count = count;
  A.total = items;
  if (items == 0) return WMPC_OK;
  if (plan->lds > 64 * 1024)
    WMPC_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_osqp_hetero),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)plan->lds));
  // persistent grid: as many wavefronts as the LDS of a CU holds, on every CU
  const int64_t per_cu = std::max<int64_t>(1, (int64_t)(kMaxLdsBytes / std::max<size_t>(plan->lds, 1)));
  const int64_t grid = std::min<int64_t>(items, 256 * per_cu);
  int32_t* counter = reinterpret_cast<int32_t*>(work);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_counter_zero, dim3(1), dim3(kWave), 0, st, counter);
  int rc = check_launch("hetero counter");
  if (rc) return rc;
  hipLaunchKernelGGL(k_osqp_hetero, dim3((unsigned)grid), dim3(kWave), plan->lds, st, plan->dev, A,
                     counter, max_iter);
  return check_launch("osqp hetero");
}

int wmpc_gather_rows_f32(const int32_t* rows, int32_t nrows, int32_t in_rows, const float* in,
                         int64_t in_ld, float* out, int64_t out_ld, int64_t batch, void* stream) {
  clear_error();
  WMPC_REQUIRE(nrows >= 0 && nrows <= 65535 && in_rows >= 0, "need 0 <= nrows <= 65535");
  WMPC_REQUIRE(batch >= 0 && in_ld >= batch && out_ld >= batch, "need 0 <= batch <= ld");
  if (batch == 0 || nrows == 0) return WMPC_OK;
  WMPC_REQUIRE(rows && in && out, "null argument");
  WMPC_REQUIRE(in != out, "out must not alias in");
  hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((batch + 255) / 256), (unsigned)nrows), dim3(256), 0,
                     (hipStream_t)stream, rows, in_rows, in, in_ld, out, out_ld, batch);
  return check_launch("gather rows");
}

}  // extern "C"
