// Fused fixed-iteration OSQP ADMM for MI355X (gfx950) behind include/warpmpc_hip.h.
//
// One launch runs, per instance (one lane each): the KKT value assembly and
// LDLᵀ factorization (solver.py:764-776, 799-804), `max_iter` iterations of
// the ADMM step of solver.py:806-820 (rhs build, quasi-definite solve,
// relaxation, box projection, dual update), the unscaling and the residual /
// objective evaluation of solver.py:863-872.  The iterate lives in the caller's
// x / z / y output buffers (batch-major), the solve in the work buffer.
#include <cstring>
#include <string>
#include <vector>

#include "wmpc_qdldl_device.h"

namespace wmpc {

constexpr int kOsqpBlock = 256;

struct OsqpDev {
  int32_t n = 0, m = 0, nnz_p = 0, nnz_a = 0, nnz_kkt = 0, nnz_p_full = 0;
  float sigma = 0.f, alpha = 0.f;
  // KKT value sources, indexed by original KKT slot
  const int32_t* src_kind = nullptr;
  const int32_t* src_idx = nullptr;
  const float* src_scale = nullptr;
  const float* src_off = nullptr;
  const float* q_scale = nullptr;
  const float* sd = nullptr;
  const float* sdinv = nullptr;
  const float* se = nullptr;
  const float* seinv = nullptr;
  const float* y_scale = nullptr;
  const float* y_unscale = nullptr;
  const float* rho = nullptr;
  const float* rho_inv = nullptr;
  const int32_t* a_rows = nullptr;
  const int32_t* a_cols = nullptr;
  const int32_t* pf_rows = nullptr;
  const int32_t* pf_cols = nullptr;
  const int32_t* pf_map = nullptr;
};

}  // namespace wmpc

struct wmpc_osqp_plan {
  wmpc::OsqpDev d;
  const wmpc_qdldl_plan* kkt = nullptr;
  void* block = nullptr;
};

namespace wmpc {

// rhs of the KKT solve in original order, built from the iterate on the fly:
// [sigma x - q_scaled; z - rho⁻¹ y]  (solver.py:808-810)
struct AdmmRhs {
  const OsqpDev* dv;
  // not __restrict__: these alias the iterate the kernel updates between solves
  const float* x;
  const float* z;
  const float* y;
  const float* q;
  int64_t ld;
  float sigma;
  int32_t n;
  __device__ __forceinline__ float operator()(int32_t o, int64_t b) const {
    if (o < n) return sigma * x[(int64_t)o * ld + b] - q[(int64_t)o * ld + b] * dv->q_scale[o];
    const int32_t i = o - n;
    return z[(int64_t)i * ld + b] - dv->rho_inv[i] * y[(int64_t)i * ld + b];
  }
};

template <bool SHARED, bool FACTOR>
__global__ __launch_bounds__(kOsqpBlock) void k_osqp_admm(
    QdldlTables t, OsqpDev dv, const float* __restrict__ pv, const float* __restrict__ av,
    int64_t pm_ld, const float* __restrict__ q, const float* __restrict__ l,
    const float* __restrict__ u, const float* __restrict__ x0, const float* __restrict__ z0,
    const float* __restrict__ y0, float* __restrict__ lx, float* __restrict__ dinv,
    int64_t f_ld, float* __restrict__ w, float* __restrict__ x, float* __restrict__ z,
    float* __restrict__ y, float* __restrict__ prim, float* __restrict__ dual,
    float* __restrict__ obj, int32_t max_iter, int64_t batch, int64_t ld) {
  extern __shared__ float ys[];
  const int64_t b = (int64_t)blockIdx.x * kOsqpBlock + threadIdx.x;
  if (b >= batch) return;
  const int64_t pb = pm_ld ? b : 0;
  const int64_t pes = pm_ld ? pm_ld : 1;  // entry stride of P/A (shared: one contiguous column)
  const int32_t n = dv.n, m = dv.m;
  if constexpr (FACTOR) {
    // per-instance KKT assembly + factorization; d goes to the (later reused) work rows
    KktSource src{pv, av, pm_ld, dv.src_kind, dv.src_idx, dv.src_scale, dv.src_off};
    factor_lane<KktSource, false>(t, src, pb, b, ld, lx, nullptr, dinv, ys + threadIdx.x,
                                  kOsqpBlock, ArrayRhs{nullptr, 0}, nullptr);
  }
  const float* __restrict__ lxr = lx;
  const float* __restrict__ dinvr = dinv;
  // scaled initial iterate (solver.py:842-847)
  for (int32_t i = 0; i < n; ++i)
    x[(int64_t)i * ld + b] = x0 ? x0[(int64_t)i * ld + b] * dv.sdinv[i] : 0.0f;
  for (int32_t i = 0; i < m; ++i) {
    z[(int64_t)i * ld + b] = z0 ? z0[(int64_t)i * ld + b] * dv.se[i] : 0.0f;
    y[(int64_t)i * ld + b] = y0 ? y0[(int64_t)i * ld + b] * dv.y_scale[i] : 0.0f;
  }
  const float alpha = dv.alpha, oma = 1.0f - dv.alpha;
  const AdmmRhs rhs{&dv, x, z, y, q, ld, dv.sigma, n};
  for (int32_t it = 0; it < max_iter; ++it) {
    forward_rows<SHARED>(t, lxr, f_ld, b, ld, rhs, w);
    backward_cols<SHARED>(t, lxr, dinvr, f_ld, b, ld, w);
    for (int32_t i = 0; i < n; ++i) {
      const int64_t xi = (int64_t)i * ld + b;
      const float xt = w[(int64_t)t.pinv[i] * ld + b];
      x[xi] = alpha * xt + oma * x[xi];
    }
    for (int32_t i = 0; i < m; ++i) {
      const int64_t zi = (int64_t)i * ld + b;
      const float rinv = dv.rho_inv[i];
      const float zo = z[zi], yo = y[zi];
      const float rz = zo - rinv * yo;
      const float zt = rz + rinv * w[(int64_t)t.pinv[n + i] * ld + b];
      const float zr = alpha * zt + oma * zo;
      const float lo = l[zi] * dv.se[i], hi = u[zi] * dv.se[i];
      const float zn = fminf(fmaxf(zr + rinv * yo, lo), hi);
      y[zi] = yo + dv.rho[i] * (zr - zn);
      z[zi] = zn;
    }
  }
  // unscale (solver.py:863-865)
  for (int32_t i = 0; i < n; ++i) x[(int64_t)i * ld + b] *= dv.sd[i];
  for (int32_t i = 0; i < m; ++i) {
    z[(int64_t)i * ld + b] *= dv.seinv[i];
    y[(int64_t)i * ld + b] *= dv.y_unscale[i];
  }
  // residuals and objective with the unscaled P, A, q (solver.py:866-872)
  float* ax = w;                       // rows [0, m)
  float* px = w + (int64_t)m * ld;     // rows [m, m+n)
  float* aty = w + (int64_t)(m + n) * ld;  // rows [m+n, m+2n)
  for (int32_t i = 0; i < m; ++i) ax[(int64_t)i * ld + b] = 0.0f;
  for (int32_t i = 0; i < n; ++i) {
    px[(int64_t)i * ld + b] = 0.0f;
    aty[(int64_t)i * ld + b] = 0.0f;
  }
  for (int32_t j = 0; j < dv.nnz_a; ++j) {
    const float a = av[(int64_t)j * pes + pb];
    const int32_t r = dv.a_rows[j], c = dv.a_cols[j];
    ax[(int64_t)r * ld + b] += a * x[(int64_t)c * ld + b];
    aty[(int64_t)c * ld + b] += a * y[(int64_t)r * ld + b];
  }
  for (int32_t j = 0; j < dv.nnz_p_full; ++j) {
    const float pval = pv[(int64_t)dv.pf_map[j] * pes + pb];
    px[(int64_t)dv.pf_rows[j] * ld + b] += pval * x[(int64_t)dv.pf_cols[j] * ld + b];
  }
  float pr = 0.0f, du = 0.0f, quad = 0.0f, lin = 0.0f;
  for (int32_t i = 0; i < m; ++i)
    pr = fmaxf(pr, fabsf(ax[(int64_t)i * ld + b] - z[(int64_t)i * ld + b]));
  for (int32_t i = 0; i < n; ++i) {
    const int64_t ii = (int64_t)i * ld + b;
    const float xi = x[ii], pxi = px[ii], qi = q[ii];
    du = fmaxf(du, fabsf(pxi + qi + aty[ii]));
    quad += xi * pxi;
    lin += qi * xi;
  }
  prim[b] = pr;
  dual[b] = du;
  obj[b] = 0.5f * quad + lin;
}

// KKT assembly + factorization only (CompiledOSQP.factor).
__global__ __launch_bounds__(kOsqpBlock) void k_osqp_factor(QdldlTables t, KktSource src,
                                                            float* __restrict__ lx,
                                                            float* __restrict__ d,
                                                            float* __restrict__ dinv,
                                                            int64_t batch, int64_t ld) {
  extern __shared__ float ys[];
  const int64_t b = (int64_t)blockIdx.x * kOsqpBlock + threadIdx.x;
  if (b >= batch) return;
  const int64_t pb = src.pm_ld ? b : 0;
  factor_lane<KktSource, false>(t, src, pb, b, ld, lx, d, dinv, ys + threadIdx.x, kOsqpBlock,
                                ArrayRhs{nullptr, 0}, nullptr);
}

static int osqp_check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + " launch: " + hipGetErrorString(e));
    return WMPC_ERR_HIP;
  }
  return WMPC_OK;
}

}  // namespace wmpc

using namespace wmpc;

extern "C" {

int wmpc_osqp_plan_create(const wmpc_osqp_desc* ds, const wmpc_qdldl_plan* kkt,
                          wmpc_osqp_plan** out) {
  clear_error();
  WMPC_REQUIRE(ds && kkt && out, "null argument");
  *out = nullptr;
  const int32_t n = ds->n, m = ds->m;
  WMPC_REQUIRE(n > 0 && m >= 0 && kkt->t.n == n + m, "KKT plan dimension must be n + m");
  WMPC_REQUIRE(ds->nnz_kkt == kkt->t.nnz_a, "KKT nonzero count differs from the QDLDL plan");
  WMPC_REQUIRE(ds->kkt_base && ds->p_to_kkt && ds->p_diag_sigma && ds->a_to_kkt && ds->p_scale &&
                   ds->a_scale && ds->q_scale && ds->scaling_d && ds->scaling_dinv &&
                   ds->scaling_e && ds->scaling_einv && ds->y_scale && ds->y_unscale &&
                   ds->rho_vec && ds->rho_inv_vec && ds->a_rows && ds->a_cols &&
                   (ds->nnz_p_full == 0 || (ds->p_full_rows && ds->p_full_cols && ds->p_full_value_map)),
               "descriptor has a null array");
  std::vector<int32_t> kind(ds->nnz_kkt, 2), idx(ds->nnz_kkt, 0);
  std::vector<float> scale(ds->nnz_kkt, 0.f), off(ds->nnz_kkt);
  for (int32_t j = 0; j < ds->nnz_kkt; ++j) off[j] = (float)ds->kkt_base[j];
  for (int32_t j = 0; j < ds->nnz_p; ++j) {
    const int32_t s = ds->p_to_kkt[j];
    WMPC_REQUIRE(s >= 0 && s < ds->nnz_kkt && kind[s] == 2, "p_to_kkt out of range or repeated");
    kind[s] = 0;
    idx[s] = j;
    scale[s] = (float)ds->p_scale[j];
    off[s] = (float)ds->p_diag_sigma[j];
  }
  for (int32_t j = 0; j < ds->nnz_a; ++j) {
    const int32_t s = ds->a_to_kkt[j];
    WMPC_REQUIRE(s >= 0 && s < ds->nnz_kkt && kind[s] == 2, "a_to_kkt out of range or repeated");
    kind[s] = 1;
    idx[s] = j;
    scale[s] = (float)ds->a_scale[j];
    off[s] = 0.f;
  }
  for (int32_t j = 0; j < ds->nnz_a; ++j)
    WMPC_REQUIRE(ds->a_rows[j] >= 0 && ds->a_rows[j] < m && ds->a_cols[j] >= 0 && ds->a_cols[j] < n,
                 "a_rows/a_cols out of range");
  for (int32_t j = 0; j < ds->nnz_p_full; ++j)
    WMPC_REQUIRE(ds->p_full_rows[j] >= 0 && ds->p_full_rows[j] < n && ds->p_full_cols[j] >= 0 &&
                     ds->p_full_cols[j] < n && ds->p_full_value_map[j] >= 0 &&
                     ds->p_full_value_map[j] < ds->nnz_p,
                 "p_full arrays out of range");
  // one host blob: ints then floats, every sub-array 256-byte aligned
  std::vector<uint32_t> blob;
  std::vector<size_t> offs;
  auto app_i = [&](const int32_t* s, size_t c) {
    offs.push_back(blob.size());
    for (size_t i = 0; i < c; ++i) blob.push_back((uint32_t)s[i]);
    while (blob.size() % 64) blob.push_back(0);
  };
  auto app_f = [&](const float* s, size_t c) {
    offs.push_back(blob.size());
    for (size_t i = 0; i < c; ++i) {
      uint32_t v;
      std::memcpy(&v, &s[i], 4);
      blob.push_back(v);
    }
    while (blob.size() % 64) blob.push_back(0);
  };
  auto conv = [](const double* s, size_t c) {
    std::vector<float> o(c);
    for (size_t i = 0; i < c; ++i) o[i] = (float)s[i];
    return o;
  };
  app_i(kind.data(), kind.size());
  app_i(idx.data(), idx.size());
  app_f(scale.data(), scale.size());
  app_f(off.data(), off.size());
  const double* vecs[] = {ds->q_scale, ds->scaling_d, ds->scaling_dinv, ds->scaling_e, ds->scaling_einv,
                          ds->y_scale, ds->y_unscale, ds->rho_vec, ds->rho_inv_vec};
  const int32_t lens[] = {n, n, n, m, m, m, m, m, m};
  for (int i = 0; i < 9; ++i) {
    auto v = conv(vecs[i], lens[i]);
    app_f(v.data(), v.size());
  }
  app_i(ds->a_rows, ds->nnz_a);
  app_i(ds->a_cols, ds->nnz_a);
  app_i(ds->p_full_rows, ds->nnz_p_full);
  app_i(ds->p_full_cols, ds->nnz_p_full);
  app_i(ds->p_full_value_map, ds->nnz_p_full);

  auto* plan = new wmpc_osqp_plan();
  hipError_t e = hipMalloc(&plan->block, blob.size() * 4);
  if (e == hipSuccess) e = hipMemcpy(plan->block, blob.data(), blob.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("osqp plan upload: ") + hipGetErrorString(e));
    if (plan->block) (void)hipFree(plan->block);
    delete plan;
    return WMPC_ERR_HIP;
  }
  const uint32_t* base = static_cast<const uint32_t*>(plan->block);
  size_t i = 0;
  OsqpDev& d = plan->d;
  d.n = n;
  d.m = m;
  d.nnz_p = ds->nnz_p;
  d.nnz_a = ds->nnz_a;
  d.nnz_kkt = ds->nnz_kkt;
  d.nnz_p_full = ds->nnz_p_full;
  d.sigma = (float)ds->sigma;
  d.alpha = (float)ds->alpha;
  d.src_kind = (const int32_t*)(base + offs[i++]);
  d.src_idx = (const int32_t*)(base + offs[i++]);
  d.src_scale = (const float*)(base + offs[i++]);
  d.src_off = (const float*)(base + offs[i++]);
  d.q_scale = (const float*)(base + offs[i++]);
  d.sd = (const float*)(base + offs[i++]);
  d.sdinv = (const float*)(base + offs[i++]);
  d.se = (const float*)(base + offs[i++]);
  d.seinv = (const float*)(base + offs[i++]);
  d.y_scale = (const float*)(base + offs[i++]);
  d.y_unscale = (const float*)(base + offs[i++]);
  d.rho = (const float*)(base + offs[i++]);
  d.rho_inv = (const float*)(base + offs[i++]);
  d.a_rows = (const int32_t*)(base + offs[i++]);
  d.a_cols = (const int32_t*)(base + offs[i++]);
  d.pf_rows = (const int32_t*)(base + offs[i++]);
  d.pf_cols = (const int32_t*)(base + offs[i++]);
  d.pf_map = (const int32_t*)(base + offs[i++]);
  plan->kkt = kkt;
  *out = plan;
  return WMPC_OK;
}

int wmpc_osqp_plan_destroy(wmpc_osqp_plan* plan) {
  clear_error();
  if (!plan) return WMPC_OK;
  if (plan->block) WMPC_HIP_CHECK(hipFree(plan->block));
  delete plan;
  return WMPC_OK;
}

int wmpc_osqp_factor_f32(const wmpc_osqp_plan* plan, const float* p_values,
                         const float* a_values, int64_t pm_ld, float* lx, float* dinv,
                         int64_t batch, int64_t ld, void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && (batch == 0 || (p_values && a_values && lx && dinv)), "null argument");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  WMPC_REQUIRE(pm_ld >= batch || (pm_ld == 0 && batch <= 1),
               "pm_ld must be >= batch, or 0 with batch 1 (one shared matrix, one factor)");
  if (batch == 0) return WMPC_OK;
  const QdldlTables& t = plan->kkt->t;
  const OsqpDev& dv = plan->d;
  const int64_t grid = (batch + kOsqpBlock - 1) / kOsqpBlock;
  KktSource src{p_values, a_values, pm_ld, dv.src_kind, dv.src_idx, dv.src_scale, dv.src_off};
  hipLaunchKernelGGL(k_osqp_factor, dim3((unsigned)grid), dim3(kOsqpBlock),
                     (size_t)t.max_reach * kOsqpBlock * sizeof(float), (hipStream_t)stream, t, src,
                     lx, nullptr, dinv, batch, ld);
  return osqp_check_launch("osqp factor");
}

int wmpc_osqp_solve_f32(const wmpc_osqp_plan* plan, const float* p_values,
                        const float* a_values, int64_t pm_ld, const float* q, const float* l,
                        const float* u, const float* x0, const float* z0, const float* y0,
                        const float* lx, const float* dinv, int64_t f_ld, float* lx_work,
                        float* dinv_work, float* work, float* x, float* z, float* y, float* prim,
                        float* dual, float* obj, int32_t max_iter, int64_t batch, int64_t ld,
                        void* stream) {
  clear_error();
  WMPC_REQUIRE(plan && (batch == 0 || (p_values && a_values && q && l && u && work && x && z && y && prim && dual && obj)), "null argument");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  WMPC_REQUIRE(pm_ld == 0 || pm_ld >= batch, "pm_ld must be 0 (shared) or >= batch");
  WMPC_REQUIRE(max_iter >= 0, "max_iter must be nonnegative");
  if (batch == 0) return WMPC_OK;
  const QdldlTables& t = plan->kkt->t;
  const OsqpDev& dv = plan->d;
  const int64_t grid = (batch + kOsqpBlock - 1) / kOsqpBlock;
  const size_t lds = (size_t)t.max_reach * kOsqpBlock * sizeof(float);
  hipStream_t st = (hipStream_t)stream;
  if (lx || dinv) {
    WMPC_REQUIRE(lx && dinv, "lx and dinv must be given together");
    WMPC_REQUIRE(f_ld == 0 || f_ld >= batch, "f_ld must be 0 (shared) or >= batch");
    if (f_ld == 0)
      hipLaunchKernelGGL((k_osqp_admm<true, false>), dim3((unsigned)grid), dim3(kOsqpBlock), 0, st, t,
                         dv, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, (float*)lx,
                         (float*)dinv, (int64_t)0, work, x, z, y, prim, dual, obj, max_iter, batch, ld);
    else
      hipLaunchKernelGGL((k_osqp_admm<false, false>), dim3((unsigned)grid), dim3(kOsqpBlock), 0, st,
                         t, dv, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, (float*)lx,
                         (float*)dinv, f_ld, work, x, z, y, prim, dual, obj, max_iter, batch, ld);
    return osqp_check_launch("osqp admm");
  }
  WMPC_REQUIRE(lx_work && dinv_work, "lx_work/dinv_work required when no factor is given");
  if (pm_ld == 0) {
    // one shared matrix: factor it once (instance 0, ld 1), then solve the batch
    KktSource src{p_values, a_values, 0, dv.src_kind, dv.src_idx, dv.src_scale, dv.src_off};
    hipLaunchKernelGGL(k_osqp_factor, dim3(1), dim3(kOsqpBlock), lds, st, t, src, lx_work,
                       nullptr, dinv_work, (int64_t)1, (int64_t)1);
    int rc = osqp_check_launch("osqp shared factor");
    if (rc) return rc;
    hipLaunchKernelGGL((k_osqp_admm<true, false>), dim3((unsigned)grid), dim3(kOsqpBlock), 0, st, t,
                       dv, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, lx_work, dinv_work,
                       (int64_t)0, work, x, z, y, prim, dual, obj, max_iter, batch, ld);
    return osqp_check_launch("osqp admm");
  }
  hipLaunchKernelGGL((k_osqp_admm<false, true>), dim3((unsigned)grid), dim3(kOsqpBlock), lds, st, t,
                     dv, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, lx_work, dinv_work, ld,
                     work, x, z, y, prim, dual, obj, max_iter, batch, ld);
  return osqp_check_launch("osqp admm");
}

}  // extern "C"
