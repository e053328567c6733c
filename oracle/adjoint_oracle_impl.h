/*
 * CPU oracle of OSQP's adjoint derivative (body; included twice by
 * qdldl_oracle.c with REAL/SFX set).  TEST INFRASTRUCTURE ONLY.
 *
 * Restates, per instance and in the same operation order,
 * warpmpc/jax_osqp/solver.py:1035-1192:
 *   y_u = max(y, 0), y_l = -min(y, 0), lambda = [y_l[lower], y_u[upper]];
 *   G values a[g_map] * g_sign, Aeq values a[aeq_map];
 *   slacks = G x - h (scatter-add in G entry order), h = [-l[lower], u[upper]];
 *   M0 values: diag_first 1, diag_second 0, P block p[p_full_map], G block,
 *   Aeq block, G-lambda block lambda[g_rows] * g, slack diagonal, Aeq^T block;
 *   M values: M0 with diag_first 1 + eps, diag_second -eps;
 *   rhs = [-dx; -dy[lower]; -dy[upper]; -d_nu; 0], d_nu = y_eq >= 0 ? dy_eq : -dy_eq;
 *   sol = M^-1 rhs (QDLDL factor + solve), then `refine` steps
 *   sol -= M^-1 (M0 sol - rhs)  (M0 sol by the symmetric scatter of
 *   solver.py:1035-1044);
 *   r_yl / r_yu from sol (solver.py:1141-1154), ryl = -(r_yl * y_l),
 *   ryu = r_yu * y_u;
 *   get_vec: dq = rx, dl = ryl, du = -ryu; get_mat: dP, dA (solver.py:1166-1192).
 * The factor and solve are the up-looking routines of qdldl_oracle_impl.h.
 */

#ifndef ORACLE_ADJ_TYPES_DEFINED
#define ORACLE_ADJ_TYPES_DEFINED
typedef struct {
  int32_t n, m, n_l, n_u, n_eq, dim, nnz_adj, nnz_p, nnz_a, refine;
  double eps;
  const oracle_qdldl_pattern* kkt; /* QDLDL pattern of the adjoint upper triangle */
  const int32_t* adj_rows;         /* [nnz_adj] row of each upper entry (CSC order) */
  const int32_t* adj_cols;         /* [nnz_adj] column */
  const double* adj_base;          /* [nnz_adj] structural values (adjoint_upper.data) */
  const int32_t* lower;            /* [n_l] */
  const int32_t* upper;            /* [n_u] */
  const int32_t* eq;               /* [n_eq] */
  int32_t n_pf, n_g, n_e;
  const int32_t* p_full_map;       /* [n_pf] */
  const int32_t* g_rows;           /* [n_g] */
  const int32_t* g_cols;
  const int32_t* g_map;
  const double* g_sign;
  const int32_t* e_map;            /* [n_e] */
  const int32_t* diag_first_pos;   /* [half] */
  const int32_t* diag_second_pos;  /* [half] */
  const int32_t* p_pos;            /* [n_pf] */
  const int32_t* g_pos;            /* [n_g] */
  const int32_t* aeq_pos;          /* [n_e] */
  const int32_t* glambda_pos;      /* [n_g] */
  const int32_t* slack_pos;        /* [n_l + n_u] */
  const int32_t* aeq_t_pos;        /* [n_e] */
  const int32_t* p_rows;           /* [nnz_p] upper-P entries */
  const int32_t* p_cols;
  const int32_t* a_rows;           /* [nnz_a] */
  const int32_t* a_cols;
} oracle_adjoint_desc;
#endif

static REAL SFX(rmax)(REAL a, REAL b) { return (a != a || b != b) ? (REAL)NAN : (a > b ? a : b); }
static REAL SFX(rmin)(REAL a, REAL b) { return (a != a || b != b) ? (REAL)NAN : (a < b ? a : b); }

/* One instance.  Inputs are instance rows; outputs: sol [dim], ryl, ryu [m],
 * dq [n], dl [m], du [m], dp [nnz_p], da [nnz_a].  ws: 7*dim + 2*nnz_adj +
 * 4*m + n_ineq + 1 + nnz_l scratch. */
static void SFX(adjoint_one)(const oracle_adjoint_desc* ds, const REAL* p, const REAL* a,
                             const REAL* l, const REAL* u, const REAL* x, const REAL* y,
                             const REAL* dx, const REAL* dy, REAL* sol, REAL* ryl, REAL* ryu,
                             REAL* dq, REAL* dl, REAL* du, REAL* dp, REAL* da, REAL* ws) {
  const int32_t n = ds->n, m = ds->m, nl = ds->n_l, nu = ds->n_u, ne = ds->n_eq;
  const int32_t ni = nl + nu, half = n + ni + ne, dim = ds->dim, nz = ds->nnz_adj;
  const oracle_qdldl_pattern* pt = ds->kkt;
  REAL* v0 = ws;                 /* nnz_adj unperturbed values */
  REAL* v = v0 + nz;             /* nnz_adj perturbed values */
  REAL* rhs = v + nz;            /* dim */
  REAL* res = rhs + dim;         /* dim */
  REAL* corr = res + dim;        /* dim */
  REAL* work = corr + dim;       /* dim */
  REAL* dd = work + dim;         /* dim */
  REAL* dinv = dd + dim;         /* dim */
  REAL* yw = dinv + dim;         /* dim (factor scratch, zero) */
  REAL* yl = yw + dim;           /* m */
  REAL* yu = yl + m;             /* m */
  REAL* gx = yu + m;             /* ni */
  REAL* ryl_raw = gx + (ni > 0 ? ni : 1); /* m */
  REAL* ryu_raw = ryl_raw + m;   /* m */
  REAL* lx = ryu_raw + m;        /* nnz_l */
  for (int32_t i = 0; i < m; ++i) {
    yu[i] = SFX(rmax)(y[i], 0);
    yl[i] = -SFX(rmin)(y[i], 0);
  }
  /* slacks = G x - h */
  for (int32_t i = 0; i < ni; ++i) gx[i] = 0;
  for (int32_t k = 0; k < ds->n_g; ++k) {
    const REAL gv = a[ds->g_map[k]] * (REAL)ds->g_sign[k];
    gx[ds->g_rows[k]] += gv * x[ds->g_cols[k]];
  }
  for (int32_t i = 0; i < nl; ++i) gx[i] = gx[i] - (-l[ds->lower[i]]);
  for (int32_t i = 0; i < nu; ++i) gx[nl + i] = gx[nl + i] - u[ds->upper[i]];
  /* M0 values */
  for (int32_t k = 0; k < nz; ++k) v0[k] = (REAL)ds->adj_base[k];
  for (int32_t i = 0; i < half; ++i) {
    v0[ds->diag_first_pos[i]] = 1;
    v0[ds->diag_second_pos[i]] = 0;
  }
  for (int32_t k = 0; k < ds->n_pf; ++k) v0[ds->p_pos[k]] = p[ds->p_full_map[k]];
  for (int32_t k = 0; k < ds->n_g; ++k) {
    const REAL gv = a[ds->g_map[k]] * (REAL)ds->g_sign[k];
    v0[ds->g_pos[k]] = gv;
    const int32_t r = ds->g_rows[k];
    const REAL lam = r < nl ? yl[ds->lower[r]] : yu[ds->upper[r - nl]];
    v0[ds->glambda_pos[k]] = lam * gv;
  }
  for (int32_t k = 0; k < ds->n_e; ++k) {
    v0[ds->aeq_pos[k]] = a[ds->e_map[k]];
    v0[ds->aeq_t_pos[k]] = a[ds->e_map[k]];
  }
  for (int32_t i = 0; i < ni; ++i) v0[ds->slack_pos[i]] = gx[i];
  for (int32_t k = 0; k < nz; ++k) v[k] = v0[k];
  for (int32_t i = 0; i < half; ++i) {
    v[ds->diag_first_pos[i]] = (REAL)1 + (REAL)ds->eps;
    v[ds->diag_second_pos[i]] = -(REAL)ds->eps;
  }
  /* rhs */
  for (int32_t i = 0; i < dim; ++i) rhs[i] = 0;
  for (int32_t i = 0; i < n; ++i) rhs[i] = -dx[i];
  for (int32_t i = 0; i < nl; ++i) rhs[n + i] = -dy[ds->lower[i]];
  for (int32_t i = 0; i < nu; ++i) rhs[n + nl + i] = -dy[ds->upper[i]];
  for (int32_t i = 0; i < ne; ++i) {
    const int32_t e = ds->eq[i];
    const REAL dnu = y[e] >= 0 ? dy[e] : -dy[e];
    rhs[n + ni + i] = -dnu;
  }
  /* factor M, solve, refine against M0 */
  for (int32_t i = 0; i < dim; ++i) yw[i] = 0;
  for (int32_t k = 0; k < pt->nnz_l; ++k) lx[k] = 0;
  SFX(factor_one)(pt, v, lx, dd, dinv, yw);
  SFX(solve_one)(pt, lx, dinv, rhs, sol, work);
  for (int32_t it = 0; it < ds->refine; ++it) {
    for (int32_t i = 0; i < dim; ++i) res[i] = 0;
    for (int32_t k = 0; k < nz; ++k) res[ds->adj_rows[k]] += v0[k] * sol[ds->adj_cols[k]];
    for (int32_t k = 0; k < nz; ++k) {
      const int32_t r = ds->adj_rows[k], c = ds->adj_cols[k];
      if (r != c) res[c] += v0[k] * sol[r];
    }
    for (int32_t i = 0; i < dim; ++i) res[i] = res[i] - rhs[i];
    SFX(solve_one)(pt, lx, dinv, res, corr, work);
    for (int32_t i = 0; i < dim; ++i) sol[i] = sol[i] - corr[i];
  }
  /* map back */
  for (int32_t i = 0; i < m; ++i) ryl_raw[i] = ryu_raw[i] = 0;
  int32_t pos = half + n;
  for (int32_t i = 0; i < nl; ++i) ryl_raw[ds->lower[i]] = -sol[pos + i];
  pos += nl;
  for (int32_t i = 0; i < nu; ++i) ryu_raw[ds->upper[i]] = sol[pos + i];
  pos += nu;
  for (int32_t i = 0; i < ne; ++i) {
    const int32_t e = ds->eq[i];
    const REAL es = sol[pos + i], ey = y[e];
    const int pos_y = ey >= 0;
    ryu_raw[e] = pos_y ? es / ey : 0;
    ryl_raw[e] = pos_y ? 0 : -es / ey;
  }
  for (int32_t i = 0; i < m; ++i) {
    ryl[i] = -(ryl_raw[i] * yl[i]);
    ryu[i] = ryu_raw[i] * yu[i];
  }
  const REAL* rx = sol + half;
  for (int32_t i = 0; i < n; ++i) dq[i] = rx[i];
  for (int32_t i = 0; i < m; ++i) {
    dl[i] = ryl[i];
    du[i] = -ryu[i];
  }
  for (int32_t k = 0; k < ds->nnz_p; ++k) {
    const int32_t r = ds->p_rows[k], c = ds->p_cols[k];
    dp[k] = (REAL)0.5 * (rx[r] * x[c] + rx[c] * x[r]);
  }
  for (int32_t k = 0; k < ds->nnz_a; ++k) {
    const int32_t r = ds->a_rows[k], c = ds->a_cols[k];
    da[k] = (yu[r] - yl[r]) * rx[c] + (ryu[r] - ryl[r]) * x[c];
  }
}

/* Batch of instances (instance-major rows); p/a strides 0 = shared. */
int SFX(oracle_adjoint)(const oracle_adjoint_desc* ds, int64_t batch, const REAL* p, int64_t p_stride,
                        const REAL* a, int64_t a_stride, const REAL* l, const REAL* u, const REAL* x,
                        const REAL* y, const REAL* dx, const REAL* dy, REAL* sol, REAL* ryl, REAL* ryu,
                        REAL* dq, REAL* dl, REAL* du, REAL* dp, REAL* da) {
  if (!ds || !ds->kkt || batch < 0) return 1;
  const int64_t n = ds->n, m = ds->m, dim = ds->dim;
  const int64_t wsz = 7 * dim + 2 * (int64_t)ds->nnz_adj + 4 * m + (ds->n_l + ds->n_u + 1) +
                      ds->kkt->nnz_l + 1;
#pragma omp parallel
  {
    REAL* ws = (REAL*)malloc(sizeof(REAL) * (size_t)wsz);
#pragma omp for schedule(static)
    for (int64_t b = 0; b < batch; ++b)
      SFX(adjoint_one)(ds, p + b * p_stride, a + b * a_stride, l + b * m, u + b * m, x + b * n,
                       y + b * m, dx + b * n, dy + b * m, sol + b * dim, ryl + b * m, ryu + b * m,
                       dq + b * n, dl + b * m, du + b * m, dp + b * ds->nnz_p, da + b * ds->nnz_a, ws);
    free(ws);
  }
  return 0;
}
