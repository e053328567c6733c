// Fused fixed-iteration OSQP ADMM for MI355X (gfx950) behind include/warpmpc_hip.h.
//
// One launch runs, per instance (one lane each): the KKT value assembly and
// LDLᵀ factorization (solver.py:764-776, 799-804) unless a factor is given,
// `max_iter` iterations of the ADMM step of solver.py:806-820, the unscaling and
// the residual / objective evaluation of solver.py:863-872.
//
// Per iteration: the forward and backward substitutions (prefetched column
// sweeps of wmpc_stream.h) solve the KKT system into the stash; one pass over
// the entries (variables, then constraints: no branches, independent
// iterations) applies the relaxation, box projection and dual update and
// writes the next right-hand side [sigma x - q; z - rho⁻¹ y] in the sweep's
// first-touch order.  Scratch: stash and right-hand side, [n_kkt][ld] each.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "wmpc_blob.h"
#include "wmpc_osqp_device.h"


namespace wmpc {

template <bool SHARED, bool FACTOR>
__global__ __launch_bounds__(kWave) void k_osqp_admm(
    QdldlProg P, OsqpDev dv, const float* __restrict__ pv, const float* __restrict__ av,
    int64_t pm_ld, const float* __restrict__ q, const float* __restrict__ l,
    const float* __restrict__ u, const float* __restrict__ x0, const float* __restrict__ z0,
    const float* __restrict__ y0, float* __restrict__ lx, float* __restrict__ dinv, int64_t f_ld,
    float* __restrict__ w, float* __restrict__ rbuf, float* __restrict__ x, float* __restrict__ z,
    float* __restrict__ y, float* __restrict__ prim, float* __restrict__ dual,
    float* __restrict__ obj, int32_t max_iter, int64_t batch, int64_t ld, int32_t ipw) {
  extern __shared__ float lds[];
  // lanes past the batch work on the last instance (identical loads and stores)
  int64_t b0;
  const int64_t b = lane_instance(ipw, batch, &b0);
  admm_solve<SHARED, FACTOR>(P, dv, pv, av, pm_ld, q, l, u, x0, z0, y0, lx, dinv, f_ld, w, rbuf, x,
                             z, y, prim, dual, obj, max_iter, ld, b, lds, threadIdx.x);
}

// KKT assembly + factorization only (CompiledOSQP.factor).
__global__ __launch_bounds__(kWave) void k_osqp_factor(QdldlProg P, KktSource src,
                                                       float* __restrict__ lx,
                                                       float* __restrict__ dinv, int64_t batch,
                                                       int64_t ld, int32_t ipw) {
  extern __shared__ float lds[];
  // lanes past the batch factor the last instance again (identical values to
  // the same addresses): the table streams need the whole wavefront
  int64_t b0;
  const int64_t b = lane_instance(ipw, batch, &b0);
  const int64_t pb = src.pm_ld ? b : 0;
  factor_rows<false>(P, src, pb, Strided{lx, ld, b}, Strided{nullptr, 0, 0},
                     Strided{dinv, ld, b}, lds + (threadIdx.x & (P.lcols - 1)), NoRhs{},
                     Strided{nullptr, 0, 0});
}

// Factorization of one matrix shared by the whole batch by one wavefront.
// Row k of the up-looking recurrence (core.py:867-933): scatter the upper
// entries of column k into the row workspace Y, then for each reach entry c in
// processing order the updates Y[li[q]] -= L[q] Y[c] over the column's entries
// before row k -- distinct rows, one lane each -- and L[k,c] = Y[c] D⁻¹[c],
// D[k] -= Y[c] L[k,c].  Every value sees the same operations in the same order
// as in the one-lane factorization.  All tables and values live in LDS (all
// lanes first assemble the KKT values into the row and init records), so the
// loop touches no global memory until L is written out.  A step's record is
// read two steps ahead and each lane's L entry and target row one step ahead
// (L entries of a reach column before row k were written in earlier rows; at
// the start of a row they are read again, since the row just finished may have
// written them); lanes past a step's update count work on a dummy row.  The wavefront's LDS
// accesses execute in program order, so dependent steps need no barrier.
__global__ __launch_bounds__(kWave) void k_osqp_factor_shared(SharedFactorProg F, KktSource src,
                                                              float* __restrict__ lx,
                                                              float* __restrict__ dinv) {
  extern __shared__ float sm[];
  const int lane = threadIdx.x;
  const int32_t n = F.n, nl = F.nnz_l, ns = F.nsteps;
  int4* Rw = reinterpret_cast<int4*>(sm);
  int2* In = reinterpret_cast<int2*>(sm + 4 * (int64_t)n);
  int2* St = In + F.ninit;
  // row workspace and D⁻¹ interleaved, {Y[i], D⁻¹[i]} (n + 1 pairs: a dummy row
  // for idle lanes), and L interleaved, {row li[q], value L[q]} (nl + 1 pairs):
  // a step reads Y[c] with D⁻¹[c], and a lane its L row with its value, in one
  // 8-byte LDS access
  float* YD = reinterpret_cast<float*>(St + ns);
  int32_t* LL = reinterpret_cast<int32_t*>(YD + 2 * ((int64_t)n + 1));
  // staging: unrolled so that each lane keeps several global loads in flight.
  // The KKT values are resolved here: a row record keeps its diagonal value and
  // an init record its entry's value (as bits) in place of the KKT slot, so a
  // row's scatter reads each entry with one LDS access.
#pragma unroll 8
  for (int32_t i = lane; i < n; i += kWave) {
    int4 rw = F.rows[i];
    rw.w = __float_as_int(src(rw.w, 0));
    Rw[i] = rw;
  }
  for (int32_t i = lane; i <= n; i += kWave) reinterpret_cast<float2*>(YD)[i] = float2{0.f, 0.f};
#pragma unroll 8
  for (int32_t j = lane; j < F.ninit; j += kWave) {
    int2 e = F.init[j];
    e.y = __float_as_int(src(e.y, 0));
    In[j] = e;
  }
#pragma unroll 8
  for (int32_t q = lane; q <= nl; q += kWave)
    reinterpret_cast<int2*>(LL)[q] = int2{q < nl ? F.li[q] : 0, 0};
#pragma unroll 8
  for (int32_t r = lane; r < ns; r += kWave) St[r] = F.steps[r];
  __syncthreads();
  // this lane's update of a step: target row (n: the dummy row) and L value
  auto lane_update = [&](int2 rec, int32_t& row, float& lv) {
    const int32_t cnt = rec.x >> 16;
    const bool on = lane < cnt;
    const int32_t q = rec.y + (on ? lane : 0);  // <= nnz_l (a padding slot)
    const int2 e = reinterpret_cast<const int2*>(LL)[q];
    row = on ? e.x : n;
    lv = on ? __int_as_float(e.y) : 0.f;
  };
  const int32_t slast = ns > 0 ? ns - 1 : 0;
  int2 rec_a = ns > 0 ? St[0] : int2{0, 0};
  int2 rec_b = ns > 0 ? St[ns > 1 ? 1 : 0] : int2{0, 0};
  int32_t r = 0;
  for (int32_t k = 0; k < n; ++k) {
    const int4 row = Rw[k];  // {end of steps, init begin, init end, diagonal value bits}
    for (int32_t j = row.y + lane; j < row.z; j += kWave) {
      const int2 e = In[j];  // {row, value bits}
      YD[2 * e.x] = __int_as_float(e.y);
    }
    float dk = __int_as_float(row.w);
    int32_t ia = n;
    float la = 0.f;
    if (ns > 0) lane_update(rec_a, ia, la);
    for (; r < row.x; ++r) {
      const int32_t sx = __builtin_amdgcn_readfirstlane(rec_a.x);
      const int32_t q0 = __builtin_amdgcn_readfirstlane(rec_a.y);
      const int32_t c = sx & 0xffff, cnt = sx >> 16;
      // the same column index from the record's vector copy (equal on every
      // lane): its address skips the vector-to-scalar round trip
      const int32_t cv = rec_a.x & 0xffff;
      const int32_t i0 = ia;
      const float l0 = la;
      const float2 ycd = reinterpret_cast<const float2*>(YD)[cv];
      const float yc = ycd.x, dc = ycd.y;
      const float yi = YD[2 * i0];
      // issued after this step's Y reads so that the wait for those does not
      // also wait for these; step r's L write goes to column c's slot, which
      // step r + 1 (another column) does not read
      lane_update(rec_b, ia, la);  // step r + 1 (its row is k or starts afresh)
      rec_a = rec_b;
      rec_b = St[r + 2 < ns ? r + 2 : slast];
      YD[2 * i0] = yi - l0 * yc;
      if (cnt > kWave) {
        for (int32_t t = lane + kWave; t < cnt; t += kWave) {
          const int32_t q = q0 + t;
          const int2 e = reinterpret_cast<const int2*>(LL)[q];
          YD[2 * e.x] -= __int_as_float(e.y) * yc;
        }
      }
      const float lval = yc * dc;
      LL[2 * (q0 + cnt) + 1] = __float_as_int(lval);  // every lane stores the same value
      YD[2 * c] = 0.f;
      dk -= yc * lval;
    }
    const float di = 1.0f / dk;
    YD[2 * k + 1] = di;
    if (lane == 0) dinv[k] = di;
  }
  __syncthreads();
  for (int32_t q = lane; q < nl; q += kWave) lx[q] = __int_as_float(LL[2 * q + 1]);
}

static size_t factor_lds(const QdldlProg& P) {
  return (size_t)factor_lds_floats(P, false) * sizeof(float);
}

template <class K>
static int osqp_prepare_lds(K kernel, size_t bytes) {
  if (bytes > kMaxLdsBytes) {
    set_error("pattern needs more LDS than a workgroup can have");
    return WMPC_ERR_UNSUPPORTED;
  }
  if (bytes > 64 * 1024)
    WMPC_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return WMPC_OK;
}

// CSR-style gather table of (row, col, slot) triples grouped by `key`.
static void group_by(int32_t groups, const std::vector<int32_t>& key, const std::vector<int32_t>& other,
                     const std::vector<int32_t>& slot, std::vector<int32_t>& ptr, std::vector<int32_t>& ent) {
  ptr.assign(groups + 1, 0);
  for (int32_t k : key) ++ptr[k + 1];
  for (int32_t g = 0; g < groups; ++g) ptr[g + 1] += ptr[g];
  ent.assign(2 * key.size(), 0);
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (size_t j = 0; j < key.size(); ++j) {
    const int32_t at = fill[key[j]]++;
    ent[2 * (size_t)at] = other[j];
    ent[2 * (size_t)at + 1] = slot[j];
  }
}

// Key of the pattern data a generated ADMM kernel depends on: FNV-1a (64-bit)
// over the little-endian bytes of the int32 words {generator version, n, m,
// n_kkt, nnz_l}, p[n_kkt], lp[n_kkt + 1], li[nnz_l].  The generator
// (warpmpc_amd/codegen.py) puts the same key into the kernel's name.
constexpr int32_t kGenVersion = 1;

static uint64_t fnv1a_words(uint64_t h, const int32_t* w, int64_t count) {
  for (int64_t i = 0; i < count; ++i) {
    const uint32_t v = (uint32_t)w[i];
    for (int k = 0; k < 4; ++k) {
      h ^= (v >> (8 * k)) & 0xffu;
      h *= 0x100000001b3ull;
    }
  }
  return h;
}

static uint64_t osqp_pattern_key(const wmpc_osqp_plan* plan) {
  const wmpc_qdldl_pattern& pt = plan->kkt->pattern;
  const int32_t head[5] = {kGenVersion, plan->d.n, plan->d.m, pt.n, pt.nnz_l};
  uint64_t h = 0xcbf29ce484222325ull;
  h = fnv1a_words(h, head, 5);
  h = fnv1a_words(h, pt.p, pt.n);
  h = fnv1a_words(h, pt.lp, pt.n + 1);
  h = fnv1a_words(h, pt.li, pt.nnz_l);
  return h;
}

// Launch of an attached straight-line kernel (the argument list of the
// generated kernels, in order).
static int launch_specialized(const wmpc_osqp_plan* plan, const float* p_values,
                              const float* a_values, int64_t pm_ld, const float* q,
                              const float* l, const float* u, const float* x0, const float* z0,
                              const float* y0, const float* lx, const float* dinv, float* x,
                              float* z, float* y, float* prim, float* dual, float* obj,
                              int32_t max_iter, int64_t batch, int64_t ld, int32_t ipw,
                              int64_t grid, hipStream_t st) {
  OsqpDev dv = plan->d;
  const float* vtab = reinterpret_cast<const float*>(dv.vtab);
  const float* ctab = reinterpret_cast<const float*>(dv.ctab);
  void* args[] = {&dv,  &p_values, &a_values, &pm_ld, &q,    &l,    &u,    &x0,
                  &z0,  &y0,       &lx,       &dinv,  &x,    &z,    &y,    &prim,
                  &dual, &obj,     &max_iter, &batch, &ld,   &ipw,  &vtab, &ctab};
  WMPC_HIP_CHECK(hipModuleLaunchKernel((hipFunction_t)plan->spec_fn, (unsigned)grid, 1, 1, kWave,
                                       1, 1, 0, st, args, nullptr));
  return check_launch("osqp admm (generated)");
}

// Factorization of one matrix shared by the batch: by one wavefront when its
// LDS fits, else by the one-lane row-program kernel.
static int factor_shared(const wmpc_osqp_plan* plan, const KktSource& src, float* lx, float* dinv,
                         hipStream_t st) {
  const size_t lds = shared_factor_lds(plan->sf);
  if (plan->sf.n > 0 && lds <= kMaxLdsBytes) {
    int rc = osqp_prepare_lds(k_osqp_factor_shared, lds);
    if (rc) return rc;
    hipLaunchKernelGGL(k_osqp_factor_shared, dim3(1), dim3(kWave), lds, st, plan->sf, src, lx, dinv);
    return check_launch("osqp shared factor");
  }
  const QdldlProg& P = plan->kkt->prog;
  const size_t flds = factor_lds(P);
  int rc = osqp_prepare_lds(k_osqp_factor, flds);
  if (rc) return rc;
  hipLaunchKernelGGL(k_osqp_factor, dim3(1), dim3(kWave), flds, st, P, src, lx, dinv, (int64_t)1,
                     (int64_t)1, kWave);
  return check_launch("osqp shared factor");
}


}  // namespace wmpc

using namespace wmpc;

extern "C" {

int wmpc_osqp_plan_create(const wmpc_osqp_desc* ds, const wmpc_qdldl_plan* kkt,
                          wmpc_osqp_plan** out) {
  clear_error();
  WMPC_REQUIRE(ds && kkt && out, "null argument");
  *out = nullptr;
  WMPC_REQUIRE(kkt->block, "host-only KKT plan: create it with wmpc_qdldl_plan_create");
  const int32_t n = ds->n, m = ds->m;
  WMPC_REQUIRE(n > 0 && m >= 0 && kkt->host.n == n + m, "KKT plan dimension must be n + m");
  WMPC_REQUIRE(ds->nnz_kkt == kkt->host.nnz_a, "KKT nonzero count differs from the QDLDL plan");
  WMPC_REQUIRE(ds->kkt_base && ds->p_to_kkt && ds->p_diag_sigma && ds->a_to_kkt && ds->p_scale &&
                   ds->a_scale && ds->q_scale && ds->scaling_d && ds->scaling_dinv &&
                   ds->scaling_e && ds->scaling_einv && ds->y_scale && ds->y_unscale &&
                   ds->rho_vec && ds->rho_inv_vec && (ds->nnz_a == 0 || (ds->a_rows && ds->a_cols)) &&
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
  std::vector<int32_t> a_rows(ds->nnz_a), a_cols(ds->nnz_a), a_slot(ds->nnz_a);
  for (int32_t j = 0; j < ds->nnz_a; ++j) {
    WMPC_REQUIRE(ds->a_rows[j] >= 0 && ds->a_rows[j] < m && ds->a_cols[j] >= 0 && ds->a_cols[j] < n,
                 "a_rows/a_cols out of range");
    a_rows[j] = ds->a_rows[j];
    a_cols[j] = ds->a_cols[j];
    a_slot[j] = j;
  }
  std::vector<int32_t> pf_rows(ds->nnz_p_full), pf_cols(ds->nnz_p_full), pf_map(ds->nnz_p_full);
  for (int32_t j = 0; j < ds->nnz_p_full; ++j) {
    WMPC_REQUIRE(ds->p_full_rows[j] >= 0 && ds->p_full_rows[j] < n && ds->p_full_cols[j] >= 0 &&
                     ds->p_full_cols[j] < n && ds->p_full_value_map[j] >= 0 &&
                     ds->p_full_value_map[j] < ds->nnz_p,
                 "p_full arrays out of range");
    pf_rows[j] = ds->p_full_rows[j];
    pf_cols[j] = ds->p_full_cols[j];
    pf_map[j] = ds->p_full_value_map[j];
  }
  std::vector<int32_t> arow_ptr, arow, acol_ptr, acol, prow_ptr, prow;
  group_by(m, a_rows, a_cols, a_slot, arow_ptr, arow);
  group_by(n, a_cols, a_rows, a_slot, acol_ptr, acol);
  group_by(n, pf_rows, pf_cols, pf_map, prow_ptr, prow);
  std::vector<int32_t> vtab, ctab;
  for (int32_t c = 0; c < n + m; ++c) {
    const int32_t o = kkt->host.p[c];
    const int32_t ft = kkt->host.ftpos[c];
    if (o < n) {
      vtab.insert(vtab.end(), {c, o, ft, float_bits(ds->q_scale[o])});
    } else {
      const int32_t i = o - n;
      ctab.insert(ctab.end(), {c, i, ft, float_bits(ds->rho_inv_vec[i]), float_bits(ds->rho_vec[i]),
                               float_bits(ds->scaling_e[i]), 0, 0});
    }
  }
  // one-wavefront shared factorization (steps pack c and the update count in
  // 16 bits each: larger patterns keep the one-lane kernel)
  const wmpc_qdldl_pattern& pt = kkt->pattern;
  const int32_t nk = pt.n, nproc = pt.proc_ptr[nk], ninit = pt.init_ptr[nk];
  std::vector<int32_t> sf_rows(4 * (size_t)nk), sf_steps(2 * (size_t)std::max(nproc, 1), 0),
      sf_init(2 * (size_t)std::max(ninit, 1), 0), sf_li(pt.li, pt.li + std::max(pt.nnz_l, 1));
  bool sf_ok = nk < 65536;
  for (int32_t k = 0; k < nk; ++k) {
    sf_rows[4 * (size_t)k] = pt.proc_ptr[k + 1];
    sf_rows[4 * (size_t)k + 1] = pt.init_ptr[k];
    sf_rows[4 * (size_t)k + 2] = pt.init_ptr[k + 1];
    sf_rows[4 * (size_t)k + 3] = pt.perm_to_orig[pt.diag_pos[k]];
  }
  for (int32_t r = 0; r < nproc; ++r) {
    sf_ok = sf_ok && pt.proc_off[r] < 32768;
    sf_steps[2 * (size_t)r] = pt.proc_cols[r] | (pt.proc_off[r] << 16);
    sf_steps[2 * (size_t)r + 1] = pt.lp[pt.proc_cols[r]];
  }
  for (int32_t j = 0; j < ninit; ++j) {
    sf_init[2 * (size_t)j] = pt.init_rows[j];
    sf_init[2 * (size_t)j + 1] = pt.perm_to_orig[pt.init_pos[j]];
  }
  Blob blob;
  const size_t t_sfr = blob.add(sf_rows), t_sfs = blob.add(sf_steps), t_sfl = blob.add(sf_li),
               t_sfi = blob.add(sf_init);
  const size_t t_kind = blob.add(kind), t_idx = blob.add(idx), t_scale = blob.add(scale),
               t_off = blob.add(off), t_vtab = blob.add(vtab), t_ctab = blob.add(ctab);
  const size_t t_d = blob.add_f32(ds->scaling_d, n), t_dinv = blob.add_f32(ds->scaling_dinv, n),
               t_e = blob.add_f32(ds->scaling_e, m), t_einv = blob.add_f32(ds->scaling_einv, m),
               t_ys = blob.add_f32(ds->y_scale, m), t_yu = blob.add_f32(ds->y_unscale, m);
  const size_t t_arp = blob.add(arow_ptr), t_ar = blob.add(arow), t_acp = blob.add(acol_ptr),
               t_ac = blob.add(acol), t_prp = blob.add(prow_ptr), t_pr = blob.add(prow);
  auto* plan = new wmpc_osqp_plan();
  if (int rc = blob.upload(&plan->block, "osqp plan")) {
    delete plan;
    return rc;
  }
  OsqpDev& d = plan->d;
  d.n = n;
  d.m = m;
  d.nnz_p = ds->nnz_p;
  d.nnz_a = ds->nnz_a;
  d.nnz_kkt = ds->nnz_kkt;
  d.sigma = (float)ds->sigma;
  d.alpha = (float)ds->alpha;
  d.src_kind = blob.at<int32_t>(t_kind);
  d.src_idx = blob.at<int32_t>(t_idx);
  d.src_scale = blob.at<float>(t_scale);
  d.src_off = blob.at<float>(t_off);
  d.nvar = n;
  d.ncon = m;
  d.vtab = blob.at<int4>(t_vtab);
  d.ctab = blob.at<int4>(t_ctab);
  d.sd = blob.at<float>(t_d);
  d.sdinv = blob.at<float>(t_dinv);
  d.se = blob.at<float>(t_e);
  d.seinv = blob.at<float>(t_einv);
  d.y_scale = blob.at<float>(t_ys);
  d.y_unscale = blob.at<float>(t_yu);
  d.arow_ptr = blob.at<int32_t>(t_arp);
  d.arow = blob.at<int2>(t_ar);
  d.acol_ptr = blob.at<int32_t>(t_acp);
  d.acol = blob.at<int2>(t_ac);
  d.prow_ptr = blob.at<int32_t>(t_prp);
  d.prow = blob.at<int2>(t_pr);
  SharedFactorProg& sf = plan->sf;
  sf.n = nk;
  sf.nnz_l = pt.nnz_l;
  sf.nsteps = nproc;
  sf.nnz_kkt = ds->nnz_kkt;
  sf.ninit = ninit;
  sf.rows = blob.at<int4>(t_sfr);
  sf.steps = blob.at<int2>(t_sfs);
  sf.li = blob.at<int32_t>(t_sfl);
  sf.init = blob.at<int2>(t_sfi);
  if (!sf_ok) sf.n = 0;  // marks the one-lane fallback
  plan->kkt = kkt;
  *out = plan;
  return WMPC_OK;
}

int wmpc_osqp_plan_destroy(wmpc_osqp_plan* plan) {
  clear_error();
  if (!plan) return WMPC_OK;
  if (plan->spec_module) WMPC_HIP_CHECK(hipModuleUnload((hipModule_t)plan->spec_module));
  if (plan->block) WMPC_HIP_CHECK(hipFree(plan->block));
  delete plan;
  return WMPC_OK;
}

int64_t wmpc_osqp_work_floats(const wmpc_osqp_plan* plan, int32_t factor_mode, int64_t ld) {
  if (!plan || ld < 0 || factor_mode < 0 || factor_mode > 2) return -1;
  const int64_t nk = plan->kkt->host.n, nl = plan->kkt->host.nnz_l;
  int64_t cols = factor_mode == 2 ? ld : (factor_mode == 1 ? 1 : 0);
  return 2 * nk * ld + (nl + nk) * cols;
}

int wmpc_osqp_pattern_key(const wmpc_osqp_plan* plan, uint64_t* key) {
  clear_error();
  WMPC_REQUIRE(plan && key, "null argument");
  *key = osqp_pattern_key(plan);
  return WMPC_OK;
}

int wmpc_osqp_plan_specialize(wmpc_osqp_plan* plan, const void* code_object, int64_t size) {
  clear_error();
  WMPC_REQUIRE(plan, "null argument");
  WMPC_REQUIRE((code_object == nullptr) == (size == 0) && size >= 0,
               "code_object and size must be given together (null and 0 detach)");
  if (!code_object) {
    if (plan->spec_module) {
      WMPC_HIP_CHECK(hipModuleUnload((hipModule_t)plan->spec_module));
      plan->spec_module = nullptr;
      plan->spec_fn = nullptr;
    }
    return WMPC_OK;
  }
  // the new image replaces an attached one only once its kernel is found
  hipModule_t mod = nullptr;
  WMPC_HIP_CHECK(hipModuleLoadData(&mod, code_object));
  char name[64];
  std::snprintf(name, sizeof(name), "wmpc_admm_shared_%016llx",
                (unsigned long long)osqp_pattern_key(plan));
  hipFunction_t fn = nullptr;
  if (hipModuleGetFunction(&fn, mod, name) != hipSuccess || !fn) {
    (void)hipGetLastError();
    (void)hipModuleUnload(mod);
    set_error(std::string("code object has no kernel ") + name +
              ": it was generated for another pattern (or generator version)");
    return WMPC_ERR_INVALID;
  }
  if (plan->spec_module) {
    hipError_t e = hipModuleUnload((hipModule_t)plan->spec_module);
    plan->spec_module = mod;
    plan->spec_fn = fn;
    WMPC_HIP_CHECK(e);
    return WMPC_OK;
  }
  plan->spec_module = mod;
  plan->spec_fn = fn;
  return WMPC_OK;
}

int wmpc_osqp_plan_specialized(const wmpc_osqp_plan* plan) {
  clear_error();
  WMPC_REQUIRE(plan, "null argument");
  return plan->spec_fn ? 1 : 0;
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
  const QdldlProg& P = plan->kkt->prog;
  const OsqpDev& dv = plan->d;
  KktSource src{p_values, a_values, pm_ld, dv.src_kind, dv.src_idx, dv.src_scale, dv.src_off};
  if (pm_ld == 0 && ld == 1) return factor_shared(plan, src, lx, dinv, (hipStream_t)stream);
  const size_t lds = factor_lds(P);
  int rc = osqp_prepare_lds(k_osqp_factor, lds);
  if (rc) return rc;
  const int32_t ipw = instances_per_wave(batch, P);
  hipLaunchKernelGGL(k_osqp_factor, dim3((unsigned)wave_grid(batch, ipw)), dim3(kWave), lds,
                     (hipStream_t)stream, P, src, lx, dinv, batch, ld, ipw);
  return check_launch("osqp factor");
}

}  // extern "C"

namespace wmpc {

// Body of wmpc_osqp_solve_f32; `work_what` names the scratch in size errors.
static int osqp_solve(const wmpc_osqp_plan* plan, const float* p_values, const float* a_values,
                      int64_t pm_ld, const float* q, const float* l, const float* u,
                      const float* x0, const float* z0, const float* y0, const float* lx,
                      const float* dinv, int64_t f_ld, float* work, int64_t work_len,
                      const char* work_what, float* x, float* z, float* y, float* prim,
                      float* dual, float* obj, int32_t max_iter, int64_t batch, int64_t ld,
                      void* stream) {
  WMPC_REQUIRE(plan && (batch == 0 || (p_values && a_values && q && l && u && work && x && z && y &&
                                       prim && dual && obj)),
               "null argument");
  WMPC_REQUIRE(batch >= 0 && ld >= batch, "need 0 <= batch <= ld");
  WMPC_REQUIRE(pm_ld == 0 || pm_ld >= batch, "pm_ld must be 0 (shared) or >= batch");
  WMPC_REQUIRE(max_iter >= 0, "max_iter must be nonnegative");
  if (batch == 0) return WMPC_OK;
  WMPC_REQUIRE((lx == nullptr) == (dinv == nullptr), "lx and dinv must be given together");
  const int32_t mode = lx ? 0 : (pm_ld == 0 ? 1 : 2);
  WMPC_REQUIRE(work_len >= wmpc_osqp_work_floats(plan, mode, ld), work_what);
  const QdldlProg& P = plan->kkt->prog;
  const OsqpDev& dv = plan->d;
  const int64_t nk = P.n, nl = P.nnz_l;
  const int32_t ipw = instances_per_wave(batch, P);
  const int64_t grid = wave_grid(batch, ipw);
  hipStream_t st = (hipStream_t)stream;
  float* stash = work;
  float* rbuf = work + nk * ld;
  float* fwork = rbuf + nk * ld;
  if (mode == 0) {
    WMPC_REQUIRE(f_ld == 0 || f_ld >= batch, "f_ld must be 0 (shared) or >= batch");
    const bool sh = f_ld == 0;
    const size_t lds = admm_lds(P, sh, false);
    if (sh && plan->spec_fn)
      return launch_specialized(plan, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, lx, dinv, x,
                                z, y, prim, dual, obj, max_iter, batch, ld, ipw, grid, st);
    if (sh) {
      int rc = osqp_prepare_lds(k_osqp_admm<true, false>, lds);
      if (rc) return rc;
      hipLaunchKernelGGL((k_osqp_admm<true, false>), dim3((unsigned)grid), dim3(kWave), lds, st, P,
                         dv, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, (float*)lx,
                         (float*)dinv, (int64_t)0, stash, rbuf, x, z, y, prim, dual, obj, max_iter,
                         batch, ld, ipw);
    } else {
      int rc = osqp_prepare_lds(k_osqp_admm<false, false>, lds);
      if (rc) return rc;
      hipLaunchKernelGGL((k_osqp_admm<false, false>), dim3((unsigned)grid), dim3(kWave), lds, st,
                         P, dv, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, (float*)lx,
                         (float*)dinv, f_ld, stash, rbuf, x, z, y, prim, dual, obj, max_iter, batch,
                         ld, ipw);
    }
    return check_launch("osqp admm");
  }
  if (mode == 1) {
    // one shared matrix: factor it once (one lane), then solve the batch
    float* lxw = fwork;
    float* dinvw = lxw + nl;
    KktSource src{p_values, a_values, 0, dv.src_kind, dv.src_idx, dv.src_scale, dv.src_off};
    int rc = factor_shared(plan, src, lxw, dinvw, st);
    if (rc) return rc;
    if (plan->spec_fn)
      return launch_specialized(plan, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, lxw, dinvw, x,
                                z, y, prim, dual, obj, max_iter, batch, ld, ipw, grid, st);
    const size_t lds = admm_lds(P, true, false);
    rc = osqp_prepare_lds(k_osqp_admm<true, false>, lds);
    if (rc) return rc;
    hipLaunchKernelGGL((k_osqp_admm<true, false>), dim3((unsigned)grid), dim3(kWave), lds, st, P,
                       dv, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, lxw, dinvw, (int64_t)0,
                       stash, rbuf, x, z, y, prim, dual, obj, max_iter, batch, ld, ipw);
    return check_launch("osqp admm");
  }
  float* lxw = fwork;
  float* dinvw = lxw + nl * ld;
  const size_t lds = admm_lds(P, false, true);
  int rc = osqp_prepare_lds(k_osqp_admm<false, true>, lds);
  if (rc) return rc;
  hipLaunchKernelGGL((k_osqp_admm<false, true>), dim3((unsigned)grid), dim3(kWave), lds, st, P, dv,
                     p_values, a_values, pm_ld, q, l, u, x0, z0, y0, lxw, dinvw, ld, stash, rbuf, x,
                     z, y, prim, dual, obj, max_iter, batch, ld, ipw);
  return check_launch("osqp admm");
}

}  // namespace wmpc

// Scratch of wmpc_osqp_solve_f32_noalloc: a fixed device array, one copy per
// GPU the library runs on, held until the process ends.  Host code never uses
// this name as an address (on the host it is only a stand-in for the symbol):
// the device address comes from hipGetSymbolAddress, once per device.
static constexpr int64_t kNoAllocScratchFloats = int64_t(16) << 20;  // 64 MiB
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace wmpc {

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the current device (resolved on first use).
static int scratch_of_current_device(float** out) {
  int dev = 0;
  WMPC_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices) {
    set_error("device index beyond the scratch table of the no-allocation entry point");
    return WMPC_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> guard(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    WMPC_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<float*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return WMPC_OK;
}

}  // namespace wmpc

extern "C" {

int wmpc_osqp_solve_f32(const wmpc_osqp_plan* plan, const float* p_values,
                        const float* a_values, int64_t pm_ld, const float* q, const float* l,
                        const float* u, const float* x0, const float* z0, const float* y0,
                        const float* lx, const float* dinv, int64_t f_ld, float* work,
                        int64_t work_len, float* x, float* z, float* y, float* prim, float* dual,
                        float* obj, int32_t max_iter, int64_t batch, int64_t ld, void* stream) {
  clear_error();
  return osqp_solve(plan, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, lx, dinv, f_ld, work,
                    work_len, "work is smaller than wmpc_osqp_work_floats(plan, mode, ld)", x, z, y,
                    prim, dual, obj, max_iter, batch, ld, stream);
}

int64_t wmpc_osqp_noalloc_max_floats(void) { return kNoAllocScratchFloats; }

int wmpc_osqp_solve_f32_noalloc(const wmpc_osqp_plan* plan, const float* p_values,
                                const float* a_values, int64_t pm_ld, const float* q,
                                const float* l, const float* u, const float* x0, const float* z0,
                                const float* y0, const float* lx, const float* dinv, int64_t f_ld,
                                float* x, float* z, float* y, float* prim, float* dual, float* obj,
                                int32_t max_iter, int64_t batch, int64_t ld, void* stream) {
  clear_error();
  if (batch == 0) {
    WMPC_REQUIRE(plan, "null argument");
    return WMPC_OK;
  }
  float* work = nullptr;
  if (int rc = scratch_of_current_device(&work)) return rc;
  return osqp_solve(plan, p_values, a_values, pm_ld, q, l, u, x0, z0, y0, lx, dinv, f_ld, work,
                    kNoAllocScratchFloats,
                    "batch too large for the no-allocation entry point: "
                    "wmpc_osqp_work_floats(plan, mode, ld) exceeds wmpc_osqp_noalloc_max_floats()",
                    x, z, y, prim, dual, obj, max_iter, batch, ld, stream);
}

}  // extern "C"
