// Batched fixed-pattern QDLDL factor / solve for MI355X (gfx950) behind the C ABI
// of include/warpmpc_hip.h.
//
// Replaces the Warp kernels of warpmpc/jax_qdldl/warp_backend.py:41-454 and the
// launch glue of :583-1076.  Launch shape: one 64-lane wavefront per workgroup,
// one lane per instance, batch-major data, the working values of the current
// rows in LDS (see wmpc_program.h), each factor/solve one launch.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wmpc_plan.h"
#include "wmpc_stream.h"


namespace wmpc {

static_assert(kRecChunk == kChunkRecs && kRecMark == kMarkBit && kRecNoSlot == kNoSlot &&
                  kRecSlotMask == kSlotMask && kRecVMax == kVPer && kRecL == kLPer &&
                  kFirstTouch == kFirstTouchBit,
              "host record layout must match the device sweeps");

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

static std::atomic<int32_t> g_ipw{0};
int32_t instances_per_wave(int64_t batch, const QdldlProg& P) {
  return std::min(P.ipw ? P.ipw : instances_per_wave(batch), P.lcols);
}

int32_t instances_per_wave(int64_t batch) {
  const int32_t forced = g_ipw.load(std::memory_order_relaxed);
  if (forced) return forced;
  int32_t ipw = 8;
  while (ipw < kWave && (batch + 2 * ipw - 1) / (2 * ipw) >= kFillWaves) ipw *= 2;
  return ipw;
}

// ---------------------------------------------------------------------------
// kernels

__global__ __launch_bounds__(kWave) void k_qdldl_factor(QdldlProg P, DirectSource src,
                                                        float* __restrict__ lx,
                                                        float* __restrict__ d,
                                                        float* __restrict__ dinv, int64_t batch,
                                                        int64_t ld, int32_t ipw) {
  extern __shared__ float lds[];
  // lanes past the batch factor the last instance again (identical values to
  // the same addresses): the table streams need the whole wavefront
  int64_t b0;
  const int64_t b = lane_instance(ipw, batch, &b0);
  factor_rows<false>(P, src, b, Strided{lx, ld, b}, Strided{d, ld, b}, Strided{dinv, ld, b},
                     lds + (threadIdx.x & (P.lcols - 1)), NoRhs{}, Strided{nullptr, 0, 0});
}

__global__ __launch_bounds__(kWave) void k_qdldl_factor_solve(
    QdldlProg P, DirectSource src, const float* __restrict__ rhs, float* __restrict__ x,
    float* __restrict__ lx, float* __restrict__ d, float* __restrict__ dinv,
    float* __restrict__ w, int64_t batch, int64_t ld, int32_t ipw) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  // every lane takes part in the wave-shared staging of the sweeps, lanes
  // past the batch on the last instance
  int64_t b0;
  const int64_t b = lane_instance(ipw, batch, &b0);
  const Strided stash{w, ld, b};
  factor_rows<true>(P, src, b, Strided{lx, ld, b}, Strided{d, ld, b}, Strided{dinv, ld, b},
                    lds + (lane & (P.lcols - 1)), ArrayRhs{CStrided{rhs, ld, b}}, stash);
  const SweepProg S = sweep_prog(P);
  bwd_cols<false>(S, lx, ld, b, w, ld, b, lds, lane, ArrayOut{Strided{x, ld, b}});
}

// SHARED: one factor ([nnz_l], [n]) for the whole batch.
template <bool SHARED>
__global__ __launch_bounds__(kWave) void k_qdldl_solve(QdldlProg P, const float* __restrict__ lx,
                                                       const float* __restrict__ dinv,
                                                       int64_t f_ld, const float* __restrict__ rhs,
                                                       float* __restrict__ x,
                                                       float* __restrict__ w, int64_t batch,
                                                       int64_t ld, int32_t ipw) {
  extern __shared__ float lds[];
  const int lane = threadIdx.x;
  int64_t b0;
  const int64_t b = lane_instance(ipw, batch, &b0);
  const int64_t fb = SHARED ? 0 : b, fs = SHARED ? 1 : f_ld;
  // right-hand side in first-touch order, staged in the x rows
#pragma unroll 8
  for (int32_t i = 0; i < P.n; ++i) x[(int64_t)i * ld + b] = rhs[(int64_t)P.rperm[i] * ld + b];
  const SweepProg S = sweep_prog(P);
  fwd_cols<SHARED>(S, lx, dinv, fs, fb, x, w, ld, b, lds, lane);
  bwd_cols<SHARED>(S, lx, fs, fb, w, ld, b, lds, lane, ArrayOut{Strided{x, ld, b}});
}

// ---------------------------------------------------------------------------
// plan construction (host)

static int upload_program(wmpc_qdldl_plan* plan) {
  const HostProgram& h = plan->host;
  std::vector<int32_t> blob;
  std::vector<size_t> offs;
  auto append = [&](const std::vector<int32_t>& v) {
    offs.push_back(blob.size());
    blob.insert(blob.end(), v.begin(), v.end());
    blob.push_back(0);                      // keep empty arrays addressable
    while (blob.size() % 64) blob.push_back(0);  // 256-byte aligned sub-arrays
  };
  append(h.row);
  append(h.init);
  append(h.proc);
  append(h.upd);
  append(h.bcol);
  append(h.bent);
  append(h.p);
  append(h.fcol);
  append(h.bcol2);
  append(h.fent);
  append(h.rperm);
  append(h.frec);
  append(h.brec);
  void* dev = nullptr;
  WMPC_HIP_CHECK(hipMalloc(&dev, blob.size() * sizeof(int32_t)));
  hipError_t e = hipMemcpy(dev, blob.data(), blob.size() * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(dev);
    set_error(std::string("plan upload: ") + hipGetErrorString(e));
    return WMPC_ERR_HIP;
  }
  if (plan->block) WMPC_HIP_CHECK(hipFree(plan->block));
  plan->block = dev;
  const int32_t* base = static_cast<const int32_t*>(dev);
  QdldlProg& P = plan->prog;
  P.n = h.n;
  P.nnz_a = h.nnz_a;
  P.nnz_l = h.nnz_l;
  P.ymax = h.ymax;
  P.ncol = h.ncol;
  P.nact = h.nact;
  P.nbw = h.nbw;
  P.ninit = (int32_t)(h.init.size() / 2);
  P.nproc = (int32_t)(h.proc.size() / 4);
  P.nupd = (int32_t)(h.upd.size() / 2);
  size_t i = 0;
  P.row = reinterpret_cast<const int4*>(base + offs[i++]);
  P.init = reinterpret_cast<const int2*>(base + offs[i++]);
  P.proc = reinterpret_cast<const int4*>(base + offs[i++]);
  P.upd = reinterpret_cast<const int2*>(base + offs[i++]);
  P.bcol = reinterpret_cast<const int2*>(base + offs[i++]);
  P.bent = base + offs[i++];
  P.p = base + offs[i++];
  P.fcol = reinterpret_cast<const int4*>(base + offs[i++]);
  P.bcol2 = reinterpret_cast<const int4*>(base + offs[i++]);
  P.fent = base + offs[i++];
  P.rperm = base + offs[i++];
  P.frec = base + offs[i++];
  P.brec = base + offs[i++];
  P.fchunks = h.fchunks;
  P.bchunks = h.bchunks;
  for (int j = 0; j < 3; ++j) {
    P.f1[j] = h.f1[j];
    P.b1[j] = h.b1[j];
  }
  return WMPC_OK;
}

// Derive the programs for the plan's LDS budget (bytes of factor windows per
// workgroup): live-L slots get whatever the windows of the fused factor+solve
// leave.  The windows take `lcols` LDS columns: 64 (one per lane) when the
// factor windows and the sweeps fit in the largest workgroup LDS, else the
// largest of 32, 16, 8 that fits, the lanes of one instance sharing a column.
// A default-budget plan whose factor windows exceed the budget gets the
// smallest budget (in 4 KiB steps) that holds them; an explicit budget must.
static int derive(wmpc_qdldl_plan* plan, bool upload = true, bool explicit_budget = false) {
  HostProgram probe;
  std::string err = build_program(&plan->pattern, 0, probe);
  WMPC_REQUIRE(err.empty(), err.c_str());
  const int64_t fixed = probe.ymax + 2 * (int64_t)probe.ncol + 1;  // factor windows, in slots
  const int64_t ring = (int64_t)(kLPer + 2 * kVPer) * kWave;         // sweep chunk rings, floats
  auto need_floats = [&](int64_t lc) { return std::max(fixed * lc, probe.nbw * lc + ring); };
  const int64_t limit = (int64_t)kMaxLdsBytes / 4;
  int32_t lc = kWave;
  while (lc > 8 && need_floats(lc) > limit) lc /= 2;
  WMPC_REQUIRE(need_floats(lc) <= limit,
               "the row windows of this pattern exceed a workgroup's LDS even at 8 instances "
               "per workgroup");
  if (!explicit_budget) {
    const size_t need = ((size_t)(fixed * lc) * sizeof(float) + 4095) / 4096 * 4096;
    if (need > plan->lds_budget) plan->lds_budget = std::min(need, kMaxLdsBytes);
  }
  const int64_t slots = (int64_t)(plan->lds_budget / (lc * sizeof(float)));
  WMPC_REQUIRE(slots >= fixed, "LDS budget too small for the row windows of this pattern");
  int32_t cap = (int32_t)std::min<int64_t>(slots - fixed, INT32_MAX);
  err = build_program(&plan->pattern, cap, plan->host);
  WMPC_REQUIRE(err.empty(), err.c_str());
  if (!upload) {
    plan->prog = QdldlProg();
    plan->prog.n = plan->host.n;
    plan->prog.ymax = plan->host.ymax;
    plan->prog.ncol = plan->host.ncol;
    plan->prog.nact = plan->host.nact;
    plan->prog.nbw = plan->host.nbw;
    plan->prog.nnz_l = plan->host.nnz_l;
    plan->prog.lcols = lc;
    plan->prog.ipw = plan->ipw;
    return WMPC_OK;
  }
  const int rc = upload_program(plan);
  plan->prog.lcols = lc;
  plan->prog.ipw = plan->ipw;
  ++plan->generation;
  return rc;
}

// LDS bytes of the kernels
static size_t lds_bytes_factor(const QdldlProg& P, bool fwd) {
  int64_t f = factor_lds_floats(P, fwd);
  if (fwd) f = std::max(f, sweep_lds_floats(sweep_prog(P), false));
  return (size_t)f * sizeof(float);
}
static size_t lds_bytes_solve(const QdldlProg& P, bool shared) {
  return (size_t)sweep_lds_floats(sweep_prog(P), shared) * sizeof(float);
}

template <class K>
static int prepare_lds(K kernel, size_t bytes) {
  if (bytes > kMaxLdsBytes) {
    set_error("pattern needs more LDS than a workgroup can have");
    return WMPC_ERR_UNSUPPORTED;
  }
  if (bytes > 64 * 1024)
    WMPC_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return WMPC_OK;
}

}  // namespace wmpc

using namespace wmpc;

extern "C" {

const char* wmpc_last_error(void) { return g_last_error.c_str(); }
const char* wmpc_version(void) { return "warpmpc_hip 0.2.0 gfx950"; }

static int plan_create(const wmpc_qdldl_pattern* pattern, wmpc_qdldl_plan** out, bool upload) {
  clear_error();
  WMPC_REQUIRE(pattern && out, "null argument");
  *out = nullptr;
  WMPC_REQUIRE(pattern->n > 0 && pattern->nnz_a > 0 && pattern->nnz_l >= 0,
               "pattern sizes must be positive");
  const int32_t* ptrs[] = {pattern->perm_to_orig, pattern->diag_pos, pattern->init_ptr,
                           pattern->init_rows,    pattern->init_pos, pattern->proc_ptr,
                           pattern->proc_cols,    pattern->proc_lidx, pattern->proc_off,
                           pattern->lp,           pattern->li,       pattern->p,
                           pattern->pinv};
  const int64_t n = pattern->n;
  WMPC_REQUIRE(pattern->init_ptr && pattern->proc_ptr, "pattern has a null table");
  const int64_t n_init = pattern->init_ptr[n], n_proc = pattern->proc_ptr[n];
  WMPC_REQUIRE(n_init >= 0 && n_proc >= 0, "init_ptr/proc_ptr must be nonnegative");
  const int64_t lens[] = {pattern->nnz_a, n, n + 1, n_init, n_init, n + 1, n_proc, n_proc,
                          n_proc, n + 1, pattern->nnz_l, n, n};
  // an empty table may be passed as NULL
  for (int i = 0; i < 13; ++i) WMPC_REQUIRE(ptrs[i] || lens[i] == 0, "pattern has a null table");
  auto* plan = new wmpc_qdldl_plan();
  plan->arrays.resize(13);
  for (int i = 0; i < 13; ++i) {
    if (lens[i] > 0) plan->arrays[i].assign(ptrs[i], ptrs[i] + lens[i]);
    plan->arrays[i].push_back(0);  // keeps the copy's data pointer valid when empty
  }
  wmpc_qdldl_pattern& pt = plan->pattern;
  pt.n = pattern->n;
  pt.nnz_a = pattern->nnz_a;
  pt.nnz_l = pattern->nnz_l;
  const int32_t** dst[] = {&pt.perm_to_orig, &pt.diag_pos, &pt.init_ptr, &pt.init_rows,
                           &pt.init_pos,     &pt.proc_ptr, &pt.proc_cols, &pt.proc_lidx,
                           &pt.proc_off,     &pt.lp,       &pt.li,        &pt.p,
                           &pt.pinv};
  for (int i = 0; i < 13; ++i) *dst[i] = plan->arrays[i].data();
  int rc = WMPC_OK;
  if (upload && hipGetDevice(&plan->device) != hipSuccess) {
    set_error("no HIP device");
    rc = WMPC_ERR_HIP;
  }
  if (rc == WMPC_OK) rc = derive(plan, upload);
  if (rc == WMPC_OK && upload) rc = raw_upload(plan->pattern, &plan->raw);
  if (rc != WMPC_OK) {
    if (plan->block) (void)hipFree(plan->block);
    raw_destroy(plan->raw);
    delete plan;
    return rc;
  }
  *out = plan;
  return WMPC_OK;
}

int wmpc_qdldl_plan_create(const wmpc_qdldl_pattern* pattern, wmpc_qdldl_plan** out) {
  return plan_create(pattern, out, true);
}

int wmpc_qdldl_plan_create_host(const wmpc_qdldl_pattern* pattern, wmpc_qdldl_plan** out) {
  return plan_create(pattern, out, false);
}

int wmpc_qdldl_plan_destroy(wmpc_qdldl_plan* plan) {
  clear_error();
  if (!plan) return WMPC_OK;
  if (plan->block) WMPC_HIP_CHECK(hipFree(plan->block));
  raw_destroy(plan->raw);
  delete plan;
  return WMPC_OK;
}

int wmpc_qdldl_plan_set_lds_budget(wmpc_qdldl_plan* plan, int64_t bytes) {
  clear_error();
  WMPC_REQUIRE(plan, "null plan");
  WMPC_REQUIRE(bytes > 0 && bytes <= (int64_t)kMaxLdsBytes, "LDS budget must be in (0, 160 KiB]");
  const size_t old = plan->lds_budget;
  plan->lds_budget = (size_t)bytes;
  int rc = derive(plan, plan->block != nullptr, true);
  if (rc != WMPC_OK) plan->lds_budget = old;
  return rc;
}

int wmpc_qdldl_plan_info(const wmpc_qdldl_plan* plan, int32_t* n, int32_t* nnz_a, int32_t* nnz_l,
                         int32_t* max_reach) {
  clear_error();
  WMPC_REQUIRE(plan, "null plan");
  if (n) *n = plan->host.n;
  if (nnz_a) *nnz_a = plan->host.nnz_a;
  if (nnz_l) *nnz_l = plan->host.nnz_l;
  if (max_reach) *max_reach = plan->host.ymax;
  return WMPC_OK;
}

int wmpc_qdldl_plan_slots(const wmpc_qdldl_plan* plan, int64_t* out, int32_t len) {
  clear_error();
  WMPC_REQUIRE(plan && out && len >= 9, "need a plan and 9 output entries");
  const HostProgram& h = plan->host;
  out[0] = h.ymax;
  out[1] = h.ncol;
  out[2] = h.nact;
  out[3] = h.nbw;
  out[4] = h.act_need;
  out[5] = h.global_rereads;
  out[6] = (int64_t)lds_bytes_factor(plan->prog, true);
  out[7] = (int64_t)lds_bytes_solve(plan->prog, false);
  out[8] = (int64_t)plan->lds_budget;
  if (len >= 10) out[9] = plan->prog.lcols;
  return WMPC_OK;
}

int wmpc_qdldl_plan_program(const wmpc_qdldl_plan* plan, int32_t which, int32_t* out,
                            int64_t cap, int64_t* len) {
  clear_error();
  WMPC_REQUIRE(plan && len, "null argument");
  const HostProgram& h = plan->host;
  const std::vector<int32_t>* tabs[] = {&h.row,  &h.init, &h.proc,  &h.upd,  &h.bcol,  &h.bent,
                                        &h.fcol, &h.bcol2, &h.fent, &h.ftpos, &h.frec, &h.brec};
  WMPC_REQUIRE(which >= 0 && which < 12, "table index must be in [0, 12)");
  const std::vector<int32_t>& t = *tabs[which];
  *len = (int64_t)t.size();
  if (out) {
    WMPC_REQUIRE(cap >= (int64_t)t.size(), "output buffer too small");
    std::copy(t.begin(), t.end(), out);
  }
  return WMPC_OK;
}

int wmpc_qdldl_factor_f32(const wmpc_qdldl_plan* plan, const float* a_values, float* lx, float* d,
                          float* dinv, int64_t batch, int64_t ld, void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && (batch == 0 || (a_values && lx && d && dinv)), "null argument");
  WMPC_REQUIRE(plan->block, "host-only plan: create it with wmpc_qdldl_plan_create");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  if (batch == 0) return WMPC_OK;
  const size_t lds = lds_bytes_factor(plan->prog, false);
  int rc = prepare_lds(k_qdldl_factor, lds);
  if (rc) return rc;
  const int32_t ipw = instances_per_wave(batch, plan->prog);
  hipLaunchKernelGGL(k_qdldl_factor, dim3((unsigned)wave_grid(batch, ipw)), dim3(kWave), lds,
                     (hipStream_t)stream, plan->prog, DirectSource{a_values, ld}, lx, d, dinv, batch,
                     ld, ipw);
  return check_launch("qdldl factor");
}

int wmpc_qdldl_solve_f32(const wmpc_qdldl_plan* plan, const float* lx, int64_t lx_ld,
                         const float* dinv, int64_t d_ld, const float* rhs, float* x, float* work,
                         int64_t work_len, int64_t batch, int64_t ld, void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && (batch == 0 || (lx && dinv && rhs && x && work)), "null argument");
  WMPC_REQUIRE(plan->block, "host-only plan: create it with wmpc_qdldl_plan_create");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  WMPC_REQUIRE((lx_ld == 0 && d_ld == 0) || (lx_ld >= batch && d_ld >= batch),
               "factor strides must both be 0 (shared) or both >= batch");
  WMPC_REQUIRE(lx_ld == d_ld, "lx and dinv must share one instance stride");
  if (batch == 0) return WMPC_OK;
  WMPC_REQUIRE(work_len >= (int64_t)plan->prog.n * ld, "work must hold n * ld floats");
  const bool shared = lx_ld == 0;
  const size_t lds = lds_bytes_solve(plan->prog, shared);
  const int32_t ipw = instances_per_wave(batch, plan->prog);
  const int64_t grid = wave_grid(batch, ipw);
  WMPC_REQUIRE(x != rhs, "x must not alias rhs (x stages the right-hand side)");
  if (shared) {
    int rc = prepare_lds(k_qdldl_solve<true>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(k_qdldl_solve<true>, dim3((unsigned)grid), dim3(kWave), lds,
                       (hipStream_t)stream, plan->prog, lx, dinv, (int64_t)0, rhs, x, work, batch,
                       ld, ipw);
  } else {
    int rc = prepare_lds(k_qdldl_solve<false>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(k_qdldl_solve<false>, dim3((unsigned)grid), dim3(kWave), lds,
                       (hipStream_t)stream, plan->prog, lx, dinv, lx_ld, rhs, x, work, batch, ld,
                       ipw);
  }
  return check_launch("qdldl solve");
}

int wmpc_qdldl_factor_solve_f32(const wmpc_qdldl_plan* plan, const float* a_values,
                                const float* rhs, float* x, float* lx, float* d, float* dinv,
                                float* work, int64_t work_len, int64_t batch, int64_t ld,
                                void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && (batch == 0 || (a_values && rhs && x && lx && d && dinv && work)),
               "null argument");
  WMPC_REQUIRE(plan->block, "host-only plan: create it with wmpc_qdldl_plan_create");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  if (batch == 0) return WMPC_OK;
  WMPC_REQUIRE(work_len >= (int64_t)plan->prog.n * ld, "work must hold n * ld floats");
  const size_t lds = lds_bytes_factor(plan->prog, true);
  int rc = prepare_lds(k_qdldl_factor_solve, lds);
  if (rc) return rc;
  const int32_t ipw = instances_per_wave(batch, plan->prog);
  hipLaunchKernelGGL(k_qdldl_factor_solve, dim3((unsigned)wave_grid(batch, ipw)), dim3(kWave), lds,
                     (hipStream_t)stream, plan->prog, DirectSource{a_values, ld}, rhs, x, lx, d,
                     dinv, work, batch, ld, ipw);
  return check_launch("qdldl factor+solve");
}

int wmpc_set_instances_per_wave(int32_t ipw) {
  clear_error();
  WMPC_REQUIRE(ipw == 0 || ipw == 8 || ipw == 16 || ipw == 32 || ipw == 64,
               "instances per wavefront must be 0 (automatic), 8, 16, 32 or 64");
  g_ipw.store(ipw, std::memory_order_relaxed);
  return WMPC_OK;
}

int32_t wmpc_instances_per_wave(int64_t batch) { return batch < 0 ? -1 : instances_per_wave(batch); }

int wmpc_qdldl_plan_set_instances_per_wave(wmpc_qdldl_plan* plan, int32_t ipw) {
  clear_error();
  WMPC_REQUIRE(plan, "null plan");
  WMPC_REQUIRE(ipw == 0 || ipw == 8 || ipw == 16 || ipw == 32 || ipw == 64,
               "instances per wavefront must be 0 (process-wide choice), 8, 16, 32 or 64");
  plan->ipw = ipw;
  plan->prog.ipw = ipw;
  return WMPC_OK;
}

int32_t wmpc_qdldl_plan_instances_per_wave(const wmpc_qdldl_plan* plan, int64_t batch) {
  if (!plan || batch < 0) return -1;
  return instances_per_wave(batch, plan->prog);
}

int64_t wmpc_qdldl_plan_lds_budget(const wmpc_qdldl_plan* plan) {
  return plan ? (int64_t)plan->lds_budget : -1;
}

}  // extern "C"
