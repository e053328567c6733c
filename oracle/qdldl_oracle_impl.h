/*
 * Body of the CPU oracle, included twice by qdldl_oracle.c with REAL/SFX set.
 * TEST INFRASTRUCTURE ONLY (see qdldl_oracle.c).
 */
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORACLE_TYPES_DEFINED
#define ORACLE_TYPES_DEFINED
/* Ragged (unpadded) form of the QDLDLPlan program tables; every list keeps the
 * order of the padded tables of warpmpc/jax_qdldl/core.py:147-256. */
typedef struct {
  int32_t n, nnz_a, nnz_l;
  const int32_t* perm_to_orig; /* [nnz_a] permuted value slot -> original slot */
  const int32_t* diag_pos;     /* [n] permuted slot of A_kk */
  const int32_t* init_ptr;     /* [n+1] */
  const int32_t* init_rows;    /* off-diagonal rows of permuted column k */
  const int32_t* init_pos;     /* their permuted value slots */
  const int32_t* proc_ptr;     /* [n+1] */
  const int32_t* proc_cols;    /* reach of row k in processing order */
  const int32_t* proc_lidx;    /* slot of L[k, c] */
  const int32_t* proc_off;     /* entries column c holds before row k */
  const int32_t* lp;           /* [n+1] L column pointers */
  const int32_t* li;           /* [nnz_l] L row indices */
  const int32_t* p;            /* [n] */
  const int32_t* pinv;         /* [n] */
} oracle_qdldl_pattern;

typedef struct {
  int32_t n, m, nnz_p, nnz_a, nnz_kkt, max_iter;
  int32_t nnz_p_full;
  int32_t pad_;
  double sigma, alpha;
  const oracle_qdldl_pattern* kkt;
  const double* kkt_base;
  const int32_t* p_to_kkt;
  const double* p_diag_sigma;
  const int32_t* a_to_kkt;
  const double* p_scale;
  const double* a_scale;
  const double* q_scale;
  const double* scaling_d;
  const double* scaling_dinv;
  const double* scaling_e;
  const double* scaling_einv;
  const double* y_scale;
  const double* y_unscale;
  const double* rho_vec;
  const double* rho_inv_vec;
  const int32_t* a_rows;
  const int32_t* a_cols;
  const int32_t* p_full_rows;
  const int32_t* p_full_cols;
  const int32_t* p_full_value_map;
} oracle_osqp_desc;
#endif

/* One instance of the up-looking factorization; y is an n-vector of zeros on
 * entry and is left zero on exit (core.py:867-933). */
static void SFX(factor_one)(const oracle_qdldl_pattern* pt, const REAL* a, REAL* lx, REAL* d,
                            REAL* dinv, REAL* y) {
  const int32_t n = pt->n;
  for (int32_t k = 0; k < n; ++k) {
    REAL dk = a[pt->perm_to_orig[pt->diag_pos[k]]];
    for (int32_t r = pt->init_ptr[k]; r < pt->init_ptr[k + 1]; ++r)
      y[pt->init_rows[r]] += a[pt->perm_to_orig[pt->init_pos[r]]];
    for (int32_t r = pt->proc_ptr[k]; r < pt->proc_ptr[k + 1]; ++r) {
      const int32_t c = pt->proc_cols[r];
      const int32_t off = pt->proc_off[r];
      const REAL val = y[c];
      const int32_t base = pt->lp[c];
      for (int32_t q = 0; q < off; ++q) {
        const REAL upd = -lx[base + q] * val;
        y[pt->li[base + q]] += upd;
      }
      const REAL lval = val * dinv[c];
      lx[pt->proc_lidx[r]] = lval;
      dk -= val * lval;
    }
    d[k] = dk;
    dinv[k] = (REAL)1 / dk;
    for (int32_t r = pt->proc_ptr[k]; r < pt->proc_ptr[k + 1]; ++r) y[pt->proc_cols[r]] = 0;
  }
}

/* x = Pᵀ L⁻ᵀ D⁻¹ L⁻¹ P rhs, column-oriented sweeps of core.py:935-1028. */
static void SFX(solve_one)(const oracle_qdldl_pattern* pt, const REAL* lx, const REAL* dinv,
                           const REAL* rhs, REAL* out, REAL* work) {
  const int32_t n = pt->n;
  for (int32_t i = 0; i < n; ++i) work[i] = rhs[pt->p[i]];
  for (int32_t c = 0; c < n; ++c) {
    const REAL val = work[c];
    for (int32_t q = pt->lp[c]; q < pt->lp[c + 1]; ++q) {
      const REAL upd = -lx[q] * val;
      work[pt->li[q]] += upd;
    }
  }
  for (int32_t i = 0; i < n; ++i) work[i] *= dinv[i];
  for (int32_t c = n - 1; c >= 0; --c) {
    REAL corr = 0;
    for (int32_t q = pt->lp[c]; q < pt->lp[c + 1]; ++q) {
      const REAL prod = lx[q] * work[pt->li[q]];
      corr += prod;
    }
    work[c] -= corr;
  }
  for (int32_t i = 0; i < n; ++i) out[i] = work[pt->pinv[i]];
}

/* a: [batch][nnz_a] original upper-CSC order; lx: [batch][nnz_l]; d, dinv: [batch][n]. */
int SFX(oracle_qdldl_factor)(const oracle_qdldl_pattern* pt, int64_t batch, const REAL* a, REAL* lx,
                             REAL* d, REAL* dinv) {
  if (!pt || batch < 0) return 1;
  const int64_t n = pt->n, na = pt->nnz_a, nl = pt->nnz_l;
#pragma omp parallel
  {
    REAL* y = (REAL*)calloc((size_t)(n > 0 ? n : 1), sizeof(REAL));
#pragma omp for schedule(static)
    for (int64_t b = 0; b < batch; ++b) {
      memset(lx + b * nl, 0, sizeof(REAL) * (size_t)nl);
      SFX(factor_one)(pt, a + b * na, lx + b * nl, d + b * n, dinv + b * n, y);
    }
    free(y);
  }
  return 0;
}

/* lx_stride / d_stride = 0 reuse one factor for every right-hand side. */
int SFX(oracle_qdldl_solve)(const oracle_qdldl_pattern* pt, int64_t batch, const REAL* lx,
                            int64_t lx_stride, const REAL* dinv, int64_t d_stride, const REAL* rhs,
                            REAL* out) {
  if (!pt || batch < 0) return 1;
  const int64_t n = pt->n;
#pragma omp parallel
  {
    REAL* work = (REAL*)malloc(sizeof(REAL) * (size_t)(n > 0 ? n : 1));
#pragma omp for schedule(static)
    for (int64_t b = 0; b < batch; ++b)
      SFX(solve_one)(pt, lx + b * lx_stride, dinv + b * d_stride, rhs + b * n, out + b * n, work);
    free(work);
  }
  return 0;
}

/* Fixed-iteration OSQP (solver.py:764-872).  p_values/a_values carry a batch
 * stride (0 = one matrix shared by all instances, factored once, as the
 * reference does when P and A have batch dimension 1).  x0/z0/y0 may be NULL
 * (cold start).  Outputs: x [batch][n], z, y [batch][m], prim/dual/obj [batch]. */
int SFX(oracle_osqp_solve)(const oracle_osqp_desc* ds, int64_t batch, const REAL* p_values,
                           int64_t p_stride, const REAL* a_values, int64_t a_stride, const REAL* q,
                           const REAL* l, const REAL* u, const REAL* x0, const REAL* z0,
                           const REAL* y0, REAL* x_out, REAL* z_out, REAL* y_out, REAL* prim,
                           REAL* dual, REAL* obj) {
  if (!ds || !ds->kkt || batch < 0) return 1;
  const oracle_qdldl_pattern* pt = ds->kkt;
  const int64_t n = ds->n, m = ds->m, nk = pt->n, nl = pt->nnz_l;
  const int shared = (p_stride == 0 && a_stride == 0);
  REAL* shared_lx = NULL;
  REAL* shared_d = NULL;
  REAL* shared_dinv = NULL;
  if (shared && batch > 0) {
    shared_lx = (REAL*)calloc((size_t)(nl > 0 ? nl : 1), sizeof(REAL));
    shared_d = (REAL*)malloc(sizeof(REAL) * (size_t)nk);
    shared_dinv = (REAL*)malloc(sizeof(REAL) * (size_t)nk);
    REAL* kv = (REAL*)malloc(sizeof(REAL) * (size_t)ds->nnz_kkt);
    REAL* y = (REAL*)calloc((size_t)nk, sizeof(REAL));
    for (int32_t j = 0; j < ds->nnz_kkt; ++j) kv[j] = (REAL)ds->kkt_base[j];
    for (int32_t j = 0; j < ds->nnz_p; ++j)
      kv[ds->p_to_kkt[j]] = p_values[j] * (REAL)ds->p_scale[j] + (REAL)ds->p_diag_sigma[j];
    for (int32_t j = 0; j < ds->nnz_a; ++j) kv[ds->a_to_kkt[j]] = a_values[j] * (REAL)ds->a_scale[j];
    SFX(factor_one)(pt, kv, shared_lx, shared_d, shared_dinv, y);
    free(kv);
    free(y);
  }
  const REAL sigma = (REAL)ds->sigma, alpha = (REAL)ds->alpha;
#pragma omp parallel
  {
    REAL* kv = (REAL*)malloc(sizeof(REAL) * (size_t)(ds->nnz_kkt > 0 ? ds->nnz_kkt : 1));
    REAL* lx = shared ? NULL : (REAL*)malloc(sizeof(REAL) * (size_t)(nl > 0 ? nl : 1));
    REAL* dd = shared ? NULL : (REAL*)malloc(sizeof(REAL) * (size_t)nk);
    REAL* dinv = shared ? NULL : (REAL*)malloc(sizeof(REAL) * (size_t)nk);
    REAL* yw = (REAL*)calloc((size_t)nk, sizeof(REAL));
    REAL* work = (REAL*)malloc(sizeof(REAL) * (size_t)nk);
    REAL* rhs = (REAL*)malloc(sizeof(REAL) * (size_t)nk);
    REAL* sol = (REAL*)malloc(sizeof(REAL) * (size_t)nk);
    REAL* xs = (REAL*)malloc(sizeof(REAL) * (size_t)n);
    REAL* zs = (REAL*)malloc(sizeof(REAL) * (size_t)m);
    REAL* ys = (REAL*)malloc(sizeof(REAL) * (size_t)m);
    REAL* qs = (REAL*)malloc(sizeof(REAL) * (size_t)n);
    REAL* ls = (REAL*)malloc(sizeof(REAL) * (size_t)m);
    REAL* us = (REAL*)malloc(sizeof(REAL) * (size_t)m);
    REAL* ax = (REAL*)malloc(sizeof(REAL) * (size_t)m);
    REAL* px = (REAL*)malloc(sizeof(REAL) * (size_t)n);
    REAL* aty = (REAL*)malloc(sizeof(REAL) * (size_t)n);
#pragma omp for schedule(static)
    for (int64_t b = 0; b < batch; ++b) {
      const REAL* pv = p_values + b * p_stride;
      const REAL* av = a_values + b * a_stride;
      const REAL *lxb, *dinvb;
      if (shared) {
        lxb = shared_lx;
        dinvb = shared_dinv;
      } else {
        for (int32_t j = 0; j < ds->nnz_kkt; ++j) kv[j] = (REAL)ds->kkt_base[j];
        for (int32_t j = 0; j < ds->nnz_p; ++j)
          kv[ds->p_to_kkt[j]] = pv[j] * (REAL)ds->p_scale[j] + (REAL)ds->p_diag_sigma[j];
        for (int32_t j = 0; j < ds->nnz_a; ++j) kv[ds->a_to_kkt[j]] = av[j] * (REAL)ds->a_scale[j];
        memset(lx, 0, sizeof(REAL) * (size_t)nl);
        SFX(factor_one)(pt, kv, lx, dd, dinv, yw);
        lxb = lx;
        dinvb = dinv;
      }
      const REAL* qb = q + b * n;
      const REAL* lb = l + b * m;
      const REAL* ub = u + b * m;
      for (int64_t i = 0; i < n; ++i) {
        qs[i] = qb[i] * (REAL)ds->q_scale[i];
        xs[i] = x0 ? x0[b * n + i] * (REAL)ds->scaling_dinv[i] : (REAL)0;
      }
      for (int64_t i = 0; i < m; ++i) {
        ls[i] = lb[i] * (REAL)ds->scaling_e[i];
        us[i] = ub[i] * (REAL)ds->scaling_e[i];
        zs[i] = z0 ? z0[b * m + i] * (REAL)ds->scaling_e[i] : (REAL)0;
        ys[i] = y0 ? y0[b * m + i] * (REAL)ds->y_scale[i] : (REAL)0;
      }
      for (int32_t it = 0; it < ds->max_iter; ++it) {
        for (int64_t i = 0; i < n; ++i) rhs[i] = sigma * xs[i] - qs[i];
        for (int64_t i = 0; i < m; ++i) rhs[n + i] = zs[i] - (REAL)ds->rho_inv_vec[i] * ys[i];
        SFX(solve_one)(pt, lxb, dinvb, rhs, sol, work);
        for (int64_t i = 0; i < n; ++i) xs[i] = alpha * sol[i] + ((REAL)1 - alpha) * xs[i];
        for (int64_t i = 0; i < m; ++i) {
          const REAL rinv = (REAL)ds->rho_inv_vec[i];
          const REAL ztilde = rhs[n + i] + rinv * sol[n + i];
          const REAL zr = alpha * ztilde + ((REAL)1 - alpha) * zs[i];
          /* jnp.minimum(jnp.maximum(v, l), u) (solver.py:818): NaN in any
             operand propagates */
          REAL zn = zr + rinv * ys[i];
          zn = (zn != zn) ? zn : ((ls[i] != ls[i]) ? ls[i] : (zn > ls[i] ? zn : ls[i]));
          zn = (zn != zn) ? zn : ((us[i] != us[i]) ? us[i] : (zn < us[i] ? zn : us[i]));
          ys[i] = ys[i] + (REAL)ds->rho_vec[i] * (zr - zn);
          zs[i] = zn;
        }
      }
      REAL* xo = x_out + b * n;
      REAL* zo = z_out + b * m;
      REAL* yo = y_out + b * m;
      for (int64_t i = 0; i < n; ++i) xo[i] = xs[i] * (REAL)ds->scaling_d[i];
      for (int64_t i = 0; i < m; ++i) {
        zo[i] = zs[i] * (REAL)ds->scaling_einv[i];
        yo[i] = ys[i] * (REAL)ds->y_unscale[i];
      }
      for (int64_t i = 0; i < m; ++i) ax[i] = 0;
      for (int64_t i = 0; i < n; ++i) px[i] = aty[i] = 0;
      for (int32_t j = 0; j < ds->nnz_a; ++j) {
        ax[ds->a_rows[j]] += av[j] * xo[ds->a_cols[j]];
        aty[ds->a_cols[j]] += av[j] * yo[ds->a_rows[j]];
      }
      for (int32_t j = 0; j < ds->nnz_p_full; ++j)
        px[ds->p_full_rows[j]] += pv[ds->p_full_value_map[j]] * xo[ds->p_full_cols[j]];
      REAL pr = 0, du = 0, quad = 0, lin = 0;
      /* jnp.max(jnp.abs(.)) (solver.py:866-870): NaN propagates */
      for (int64_t i = 0; i < m; ++i) {
        const REAL r = fabs((double)(ax[i] - zo[i]));
        pr = (r != r || pr != pr) ? (REAL)NAN : (r > pr ? r : pr);
      }
      for (int64_t i = 0; i < n; ++i) {
        const REAL r = fabs((double)(px[i] + qb[i] + aty[i]));
        du = (r != r || du != du) ? (REAL)NAN : (r > du ? r : du);
        quad += xo[i] * px[i];
        lin += qb[i] * xo[i];
      }
      prim[b] = pr;
      dual[b] = du;
      obj[b] = (REAL)0.5 * quad + lin;
    }
    free(kv);
    free(lx);
    free(dd);
    free(dinv);
    free(yw);
    free(work);
    free(rhs);
    free(sol);
    free(xs);
    free(zs);
    free(ys);
    free(qs);
    free(ls);
    free(us);
    free(ax);
    free(px);
    free(aty);
  }
  free(shared_lx);
  free(shared_d);
  free(shared_dinv);
  return 0;
}
