/*
 * warpmpc_hip.h — C ABI of the MI355X (gfx950) batched fixed-pattern QP core.
 *
 * Plain C: no HIP or torch types.  Device pointers are `float*` into device
 * memory owned by the caller; `stream` is a hipStream_t passed as `void*`
 * (NULL = the default stream).  Every entry point is asynchronous on `stream`,
 * never synchronises the host, and returns WMPC_OK or an error code whose text
 * `wmpc_last_error()` gives (per thread).
 *
 * Layout ("symbolic" / batch-major, the reference's transpose_work layout,
 * warpmpc/jax_qdldl/warp_backend.py:1188-1191): an array of `width` entries per
 * instance is stored as [width][ld], instance b of entry i at `ptr[i*ld + b]`,
 * with ld >= batch.  Each nonzero of K, L and D is therefore one contiguous
 * stream across instances.
 *
 * Numerical failures (zero pivots) are not errors: they produce inf/NaN in the
 * outputs exactly as the reference kernels do (warp_backend.py:97, 169).
 */
#ifndef WARPMPC_HIP_H
#define WARPMPC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WMPC_OK 0
#define WMPC_ERR_INVALID 1 /* bad argument or inconsistent pattern */
#define WMPC_ERR_HIP 2     /* a HIP runtime call failed */
#define WMPC_ERR_ALLOC 3   /* device allocation failed */
#define WMPC_ERR_UNSUPPORTED 4

/* Text of the last error raised on the calling thread ("" if none). */
const char* wmpc_last_error(void);
/* Library version string. */
const char* wmpc_version(void);

/*
 * Ragged form of one QDLDL plan (the QDLDLPlan fields of
 * warpmpc/jax_qdldl/types.py:11-53 with the padding of the program tables
 * dropped; orders unchanged).  Host pointers; only read during plan creation.
 */
typedef struct wmpc_qdldl_pattern {
  int32_t n, nnz_a, nnz_l;
  const int32_t* perm_to_orig; /* [nnz_a] permuted value slot -> original slot */
  const int32_t* diag_pos;     /* [n] permuted slot of the diagonal of column k */
  const int32_t* init_ptr;     /* [n+1] */
  const int32_t* init_rows;    /* off-diagonal rows of permuted column k */
  const int32_t* init_pos;     /* their permuted value slots */
  const int32_t* proc_ptr;     /* [n+1] */
  const int32_t* proc_cols;    /* reach of row k, processing order */
  const int32_t* proc_lidx;    /* slot of L[k, c] */
  const int32_t* proc_off;     /* entries column c holds before row k */
  const int32_t* lp;           /* [n+1] L column pointers */
  const int32_t* li;           /* [nnz_l] L row indices */
  const int32_t* p;            /* [n] permutation */
  const int32_t* pinv;         /* [n] inverse permutation */
} wmpc_qdldl_pattern;

typedef struct wmpc_qdldl_plan wmpc_qdldl_plan; /* opaque, device-resident programs */

/* Replaces the table upload of compile_qdldl_warp (warp_backend.py:1230-1286).
 * Validates the pattern, derives the kernels' row/column programs (LDS slot
 * assignment of the working values, default LDS budget 40 KiB per 64-lane
 * workgroup) and uploads them to the current device. */
int wmpc_qdldl_plan_create(const wmpc_qdldl_pattern* pattern, wmpc_qdldl_plan** out);
/* Same derivation without a device: the plan only answers the inspection
 * calls below (slots, programs); compute entry points reject it. */
int wmpc_qdldl_plan_create_host(const wmpc_qdldl_pattern* pattern, wmpc_qdldl_plan** out);
int wmpc_qdldl_plan_destroy(wmpc_qdldl_plan* plan);
/* Re-derives the programs for another LDS budget per workgroup (bytes, at most
 * 160 KiB).  Live L entries that do not fit are re-read from global memory.
 * No counterpart in the reference: the tuning knob of this backend. */
int wmpc_qdldl_plan_set_lds_budget(wmpc_qdldl_plan* plan, int64_t bytes);
/* Pattern sizes of a plan: n, nnz_a, nnz_l, max reach width. */
int wmpc_qdldl_plan_info(const wmpc_qdldl_plan* plan, int32_t* n, int32_t* nnz_a, int32_t* nnz_l,
                         int32_t* max_reach);
/* Slot counts of a plan's programs, out[0..8] = reach slots, column-window
 * slots, live-L slots in LDS, backward-window slots, live-L slots needed to keep
 * every re-read on chip, L re-reads served from global memory, LDS bytes of the
 * fused factor+solve, LDS bytes of the solve, LDS budget; out[9] (len >= 10)
 * = LDS columns of the windows: 64, or 32/16/8 for a pattern whose windows do
 * not fit in a workgroup's LDS otherwise (launches then serve at most that many
 * instances per wavefront).  len >= 9. */
int wmpc_qdldl_plan_slots(const wmpc_qdldl_plan* plan, int64_t* out, int32_t len);
/* Host copy of one program table (which: 0 row, 1 init, 2 proc, 3 upd, 4 bcol,
 * 5 bent, 6 fcol, 7 bcol2, 8 fent, 9 ftpos, 10 frec, 11 brec); *len receives
 * its length, out
 * (capacity cap) its contents if non-NULL.  For inspection and CPU replay of
 * the programs. */
int wmpc_qdldl_plan_program(const wmpc_qdldl_plan* plan, int32_t which, int32_t* out, int64_t cap,
                            int64_t* len);

/* Instances per 64-lane wavefront of every batched kernel: 0 = automatic (64
 * once the batch gives each of the 256 CUs a wavefront, i.e. batch >= 16384;
 * below that the smallest of 8, 16, 32 that still does, the lanes of one
 * instance doing identical work), or 8, 16, 32, 64.  Process-wide.  No
 * counterpart in the reference (its Warp kernels use one thread per instance):
 * the occupancy knob of this backend.  wmpc_instances_per_wave reports the
 * value a launch of `batch` instances uses (-1 on bad input). */
int wmpc_set_instances_per_wave(int32_t ipw);
int32_t wmpc_instances_per_wave(int64_t batch);
/* The same knob for the launches of one plan (every kernel that runs on its
 * pattern: QDLDL, OSQP, adjoint): 0 = the process-wide choice above, or 8, 16,
 * 32, 64; it takes precedence over wmpc_set_instances_per_wave.  Set by the
 * autotuner (warpmpc_amd/autotune.py).  wmpc_qdldl_plan_instances_per_wave
 * reports the value a launch of `batch` instances on this plan uses, capped
 * by the plan's LDS columns (-1 on bad input); wmpc_qdldl_plan_lds_budget the
 * plan's LDS budget in bytes (-1 on NULL).  No counterpart in the reference. */
int wmpc_qdldl_plan_set_instances_per_wave(wmpc_qdldl_plan* plan, int32_t ipw);
int32_t wmpc_qdldl_plan_instances_per_wave(const wmpc_qdldl_plan* plan, int64_t batch);
int64_t wmpc_qdldl_plan_lds_budget(const wmpc_qdldl_plan* plan);

/* CompiledQDLDL.factor (warp_backend.py:1302-1349; kernel :104-173):
 * a_values [nnz_a][ld] (original upper-CSC order) -> lx [nnz_l][ld],
 * d, dinv [n][ld] (permuted order). */
int wmpc_qdldl_factor_f32(const wmpc_qdldl_plan* plan, const float* a_values, float* lx, float* d,
                          float* dinv, int64_t batch, int64_t ld, void* stream);

/* CompiledQDLDL.solve (warp_backend.py:1351-1402; kernels :176-454):
 * x = Pᵀ L⁻ᵀ D⁻¹ L⁻¹ P rhs.  rhs, x [n][ld] in original order; x must not
 * alias rhs (its rows stage the right-hand side).
 * lx_ld / d_ld are the instance strides of the factor: >= batch for
 * per-instance factors, 0 for one factor ([nnz_l], [n]) shared by the batch.
 * work: scratch of work_len >= n*ld floats. */
int wmpc_qdldl_solve_f32(const wmpc_qdldl_plan* plan, const float* lx, int64_t lx_ld,
                         const float* dinv, int64_t d_ld, const float* rhs, float* x, float* work,
                         int64_t work_len, int64_t batch, int64_t ld, void* stream);

/* CompiledQDLDL.factor_and_solve (warp_backend.py:1404-1408), one launch:
 * factor with the forward substitution fused into each row, then the backward
 * sweep.  Outputs x, lx, d, dinv; work: scratch of work_len >= n*ld floats. */
int wmpc_qdldl_factor_solve_f32(const wmpc_qdldl_plan* plan, const float* a_values,
                                const float* rhs, float* x, float* lx, float* d, float* dinv,
                                float* work, int64_t work_len, int64_t batch, int64_t ld,
                                void* stream);

/*
 * float64 variants (the Warp backend's float64 path, warp_backend.py:1211-1224,
 * with x64 enabled): same layouts as the float32 entry points, one lane per
 * instance on the plan's tables directly, the operation order of
 * core.py:867-1028.  work: n * ld doubles (wmpc_qdldl_f64_work_doubles).
 * wmpc_qdldl_solve_f64: lx_ld == d_ld, 0 = one factor shared by the batch;
 * x must not alias rhs.
 */
int64_t wmpc_qdldl_f64_work_doubles(const wmpc_qdldl_plan* plan, int64_t ld);
int wmpc_qdldl_factor_f64(const wmpc_qdldl_plan* plan, const double* a_values, double* lx, double* d,
                          double* dinv, double* work, int64_t work_len, int64_t batch, int64_t ld,
                          void* stream);
int wmpc_qdldl_solve_f64(const wmpc_qdldl_plan* plan, const double* lx, int64_t lx_ld,
                         const double* dinv, int64_t d_ld, const double* rhs, double* x,
                         double* work, int64_t work_len, int64_t batch, int64_t ld, void* stream);

/*
 * Level-scheduled solve (CompiledQDLDL.solve with level_scheduled_solve=True:
 * warp_backend.py:281-454, 970-1076; levels core.py:488-510; schedule
 * warp_backend.py:1110-1126).  lanes_per_instance (1..64, a power of two)
 * lanes share one instance: a level with more than `threshold` columns is one
 * parallel step over those lanes, a run of levels of at most `threshold`
 * columns one serial step of one lane; the instance's vector lives in LDS
 * (n * 64 / lanes_per_instance floats per workgroup).  Operation order as the
 * serial column sweeps (core.py:935-1028).  Arguments as wmpc_qdldl_solve_f32,
 * without work.  wmpc_qdldl_plan_levels: number of levels of the plan (-1 on
 * a host-only plan).
 */
int32_t wmpc_qdldl_plan_levels(const wmpc_qdldl_plan* plan);
int wmpc_qdldl_solve_levels_f32(const wmpc_qdldl_plan* plan, int32_t threshold,
                                int32_t lanes_per_instance, const float* lx, int64_t lx_ld,
                                const float* dinv, int64_t d_ld, const float* rhs, float* x,
                                int64_t batch, int64_t ld, void* stream);

/*
 * Fixed-pattern OSQP (warpmpc/jax_osqp/solver.py:685-982).  The descriptor
 * carries the FixedOSQPPlan arrays (types.py:86-132) needed on the device.
 */
typedef struct wmpc_osqp_desc {
  int32_t n, m, nnz_p, nnz_a, nnz_kkt;
  int32_t nnz_p_full;
  double sigma, alpha;
  const double* kkt_base;      /* [nnz_kkt] KKT values in original upper-CSC order */
  const int32_t* p_to_kkt;     /* [nnz_p] */
  const double* p_diag_sigma;  /* [nnz_p] */
  const int32_t* a_to_kkt;     /* [nnz_a] */
  const double* p_scale;       /* [nnz_p] */
  const double* a_scale;       /* [nnz_a] */
  const double* q_scale;       /* [n] */
  const double* scaling_d;     /* [n] */
  const double* scaling_dinv;  /* [n] */
  const double* scaling_e;     /* [m] */
  const double* scaling_einv;  /* [m] */
  const double* y_scale;       /* [m] */
  const double* y_unscale;     /* [m] */
  const double* rho_vec;       /* [m] */
  const double* rho_inv_vec;   /* [m] */
  const int32_t* a_rows;       /* [nnz_a] */
  const int32_t* a_cols;       /* [nnz_a] */
  const int32_t* p_full_rows;  /* [nnz_p_full] */
  const int32_t* p_full_cols;  /* [nnz_p_full] */
  const int32_t* p_full_value_map; /* [nnz_p_full] */
} wmpc_osqp_desc;

typedef struct wmpc_osqp_plan wmpc_osqp_plan;

/* Replaces the constant setup of compile_osqp (solver.py:739-762).  `kkt` must
 * be the plan of desc's KKT pattern; it is referenced, not copied, and must
 * outlive the OSQP plan. */
int wmpc_osqp_plan_create(const wmpc_osqp_desc* desc, const wmpc_qdldl_plan* kkt,
                          wmpc_osqp_plan** out);
int wmpc_osqp_plan_destroy(wmpc_osqp_plan* plan);

/* Straight-line kernels generated per pattern.  The reference's Warp backend
 * generates its kernels per pattern at compile time (warp_backend.py:1175-1429,
 * unrolled per-pattern loops); here a generator (warpmpc_amd/codegen.py) emits
 * the whole ADMM iteration of a shared-factor solve as straight-line HIP code,
 * compiled ahead of use into a gfx950 code object.
 *
 * wmpc_osqp_pattern_key: the 64-bit key of the pattern data such a kernel
 * depends on (n, m, the KKT permutation and the structure of L); the
 * generator names its kernel wmpc_admm_shared_<key as 16 hex digits>.
 * wmpc_osqp_plan_specialize: loads a code object image (size bytes) on the
 * current device and attaches its kernel for this plan's key; solves with a
 * shared factor (factor_mode 1, or a passed factor with f_ld == 0) then launch
 * it instead of the generic record-stream kernel.  An image without the key's
 * kernel is rejected (an attached kernel then stays attached).  code_object ==
 * NULL, size == 0 detaches.
 * wmpc_osqp_plan_specialized: 1 when a kernel is attached, 0 otherwise. */
int wmpc_osqp_pattern_key(const wmpc_osqp_plan* plan, uint64_t* key);
int wmpc_osqp_plan_specialize(wmpc_osqp_plan* plan, const void* code_object, int64_t size);
int wmpc_osqp_plan_specialized(const wmpc_osqp_plan* plan);

/* CompiledOSQP.factor (solver.py:799-804): KKT value scatter + factorization.
 * p_values [nnz_p][pm_ld], a_values [nnz_a][pm_ld]; pm_ld = 0 means one matrix
 * shared by all instances (batch must then be 1 for the factor).  Outputs
 * lx [nnz_l][ld], dinv [n_kkt][ld] (d is not returned, as in the reference). */
int wmpc_osqp_factor_f32(const wmpc_osqp_plan* plan, const float* p_values,
                         const float* a_values, int64_t pm_ld, float* lx, float* dinv,
                         int64_t batch, int64_t ld, void* stream);

/* Floats of scratch wmpc_osqp_solve_f32 needs, for factor_mode 0 (a factor is
 * passed in), 1 (one shared matrix, pm_ld == 0, factored by the call) or 2
 * (per-instance matrices factored by the call): 2*n_kkt*ld for the sweep's
 * stash and staged right-hand side, plus (nnz_l + n_kkt) * (1 | ld) for the
 * factor.  -1 on bad input. */
int64_t wmpc_osqp_work_floats(const wmpc_osqp_plan* plan, int32_t factor_mode, int64_t ld);

/*
 * CompiledOSQP.solve / solve_warm_start / solve_with_factor[_warm_start]
 * (solver.py:822-982) fused into one persistent launch: per instance the KKT
 * factor (unless `lx` is given), max_iter ADMM iterations (rhs build, solve,
 * relaxation, projection, dual update), unscaling, residuals and objective.
 *
 * p_values/a_values: [nnz][pm_ld], pm_ld = 0 when shared by the batch.
 * q [n][ld], l, u [m][ld]; x0 [n][ld], z0, y0 [m][ld] may be NULL (cold start).
 * lx/dinv: if non-NULL, a precomputed factor with instance stride f_ld
 * (0 = shared, [nnz_l], [n_kkt]) and no factorization is done.
 * work: scratch of work_len >= wmpc_osqp_work_floats(plan, mode, ld) floats.
 * Outputs: x [n][ld], z, y [m][ld], prim, dual, obj [ld].
 * The box projection propagates NaN like jnp.minimum/jnp.maximum.
 */
int wmpc_osqp_solve_f32(const wmpc_osqp_plan* plan, const float* p_values,
                        const float* a_values, int64_t pm_ld, const float* q, const float* l,
                        const float* u, const float* x0, const float* z0, const float* y0,
                        const float* lx, const float* dinv, int64_t f_ld, float* work,
                        int64_t work_len, float* x, float* z, float* y, float* prim, float* dual,
                        float* obj, int32_t max_iter, int64_t batch, int64_t ld, void* stream);

/*
 * Heterogeneous batch: several fixed-pattern OSQP problem sets (different
 * plans, each with its own batch) solved in one persistent launch.  No
 * counterpart in the reference, which solves one pattern per compiled solver;
 * each instance's arithmetic is that of wmpc_osqp_solve_f32 with a per-instance
 * factor (solver.py:822-872), so results equal one wmpc_osqp_solve_f32 call
 * per pattern.  Patterns are scheduled largest-L first over a work queue.
 */
typedef struct wmpc_osqp_hetero_plan wmpc_osqp_hetero_plan;
/* One pattern's arrays for one call, as in wmpc_osqp_solve_f32 (factor always
 * computed per instance; pm_ld 0 = P/A shared by the pattern's batch). */
typedef struct wmpc_osqp_hetero_io {
  const float* p_values;
  const float* a_values;
  int64_t pm_ld;
  const float* q;
  const float* l;
  const float* u;
  const float* x0; /* warm start: x0, z0, y0 all given or all NULL */
  const float* z0;
  const float* y0;
  float* x;
  float* z;
  float* y;
  float* prim;
  float* dual;
  float* obj;
  int64_t batch, ld;
} wmpc_osqp_hetero_io;
/* 1 to 12 OSQP plans (referenced, not copied; they must outlive it).  The
 * patterns' programs are copied: a wmpc_qdldl_plan_set_lds_budget on one of
 * their KKT plans afterwards makes wmpc_osqp_solve_hetero_f32 fail with
 * WMPC_ERR_INVALID until the heterogeneous plan is created again. */
int wmpc_osqp_hetero_plan_create(const wmpc_osqp_plan* const* plans, int32_t count,
                                 wmpc_osqp_hetero_plan** out);
int wmpc_osqp_hetero_plan_destroy(wmpc_osqp_hetero_plan* plan);
/* Floats of scratch for instance strides ld[0..count). -1 on bad input. */
int64_t wmpc_osqp_hetero_work_floats(const wmpc_osqp_hetero_plan* plan, const int64_t* ld);
/* io[k]: the arrays of plans[k]. */
int wmpc_osqp_solve_hetero_f32(const wmpc_osqp_hetero_plan* plan, const wmpc_osqp_hetero_io* io,
                               int32_t max_iter, float* work, int64_t work_len, void* stream);

/* Row gather of batch-major arrays, out[i][b] = in[rows[i]][b] for i < nrows
 * (rows: device int32 array; a row outside [0, in_rows) gives NaN): the
 * one-stage shift of a previous MPC solution used as the next warm start.  No
 * counterpart in the reference (its SQP warm start reuses the previous
 * solution unshifted, sparse_mpc.py:1084-1085).  out must not alias in. */
int wmpc_gather_rows_f32(const int32_t* rows, int32_t nrows, int32_t in_rows, const float* in,
                         int64_t in_ld, float* out, int64_t out_ld, int64_t batch, void* stream);

/*
 * wmpc_osqp_solve_f32 without a work argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * scratch is a fixed device array inside the library, one per GPU, of
 * wmpc_osqp_noalloc_max_floats() floats (64 MiB), held until the process ends;
 * a call whose wmpc_osqp_work_floats(plan, mode, ld) exceeds it fails with
 * WMPC_ERR_INVALID.  The call allocates nothing and needs no part of the
 * scratch cleared (every word it reads it wrote earlier in the same call).
 * All calls on one GPU share that GPU's array: calls that overlap on one GPU
 * (other streams, other threads) race and must be serialised by the caller.
 * Other arguments, outputs and errors as wmpc_osqp_solve_f32.
 */
int64_t wmpc_osqp_noalloc_max_floats(void);
int wmpc_osqp_solve_f32_noalloc(const wmpc_osqp_plan* plan, const float* p_values,
                                const float* a_values, int64_t pm_ld, const float* q,
                                const float* l, const float* u, const float* x0, const float* z0,
                                const float* y0, const float* lx, const float* dinv, int64_t f_ld,
                                float* x, float* z, float* y, float* prim, float* dual, float* obj,
                                int32_t max_iter, int64_t batch, int64_t ld, void* stream);

/*
 * OSQP adjoint derivative (warpmpc/jax_osqp/solver.py:1035-1192).  The
 * descriptor carries the OSQPDerivativePlan fields (types.py:21-55, built by
 * solver.py:413-571) and the A entry coordinates of the FixedOSQPPlan.  With
 * half = n + n_l + n_u + n_eq the adjoint matrix is 2*half square, stored as
 * its upper triangle in CSC form.  Host pointers; read during plan creation.
 */
typedef struct wmpc_adjoint_desc {
  int32_t n, m, n_l, n_u, n_eq;
  int32_t nnz_adj, nnz_p, nnz_a;
  int32_t n_p_full, n_g, n_aeq;
  double eps;
  const int32_t* adj_indptr;        /* [2*half + 1] adjoint_upper */
  const int32_t* adj_indices;       /* [nnz_adj] */
  const double* adj_data;           /* [nnz_adj] structural values */
  const int32_t* lower_indices;     /* [n_l] */
  const int32_t* upper_indices;     /* [n_u] */
  const int32_t* eq_indices;        /* [n_eq] */
  const int32_t* p_upper_rows;      /* [nnz_p] */
  const int32_t* p_upper_cols;      /* [nnz_p] */
  const int32_t* p_full_value_map;  /* [n_p_full] */
  const int32_t* g_rows;            /* [n_g] inequality row of each G entry */
  const int32_t* g_cols;            /* [n_g] */
  const int32_t* g_a_value_map;     /* [n_g] */
  const double* g_sign;             /* [n_g] */
  const int32_t* aeq_a_value_map;   /* [n_aeq] */
  const int32_t* diag_first_pos;    /* [half] */
  const int32_t* diag_second_pos;   /* [half] */
  const int32_t* p_block_pos;       /* [n_p_full] */
  const int32_t* g_block_pos;       /* [n_g] */
  const int32_t* aeq_block_pos;     /* [n_aeq] */
  const int32_t* glambda_block_pos; /* [n_g] */
  const int32_t* slack_diag_pos;    /* [n_l + n_u] */
  const int32_t* aeq_t_block_pos;   /* [n_aeq] */
  const int32_t* a_rows;            /* [nnz_a] */
  const int32_t* a_cols;            /* [nnz_a] */
} wmpc_adjoint_desc;

typedef struct wmpc_adjoint_plan wmpc_adjoint_plan;

/* Replaces the constant setup of the adjoint in compile_osqp
 * (solver.py:984-1029).  `adj` must be the QDLDL plan of desc's adjoint
 * pattern (OSQPDerivativePlan.qdldl_plan); it is referenced, not copied, and
 * must outlive the adjoint plan. */
int wmpc_adjoint_plan_create(const wmpc_adjoint_desc* desc, const wmpc_qdldl_plan* adj,
                             wmpc_adjoint_plan** out);
int wmpc_adjoint_plan_destroy(wmpc_adjoint_plan* plan);

/*
 * KKT form of the same adjoint: implicit differentiation of the active-set
 * KKT system on the forward KKT pattern.  No counterpart in the reference,
 * whose adjoint (solver.py:1047-1139) refines a solve of the eps-perturbed
 * 2*half-square matrix M toward M0 sol = rhs, and converges at a rate of
 * eps / (sigma_min(U)^2 + eps) per step; with U nonsingular that limit is
 * sol = [0; b], U b = r.  Substituting c = +-lambda * b_lambda on rows with a
 * positive multiplier (b_nu on equality rows) turns U b = r into a symmetric
 * quasi-definite system with the forward KKT's pattern,
 *   [[P, Aᵀ], [A, diag(d)]], d = min(s, 0)/lambda | 0 (equality) | -1e30
 *   (no positive multiplier: decoupled),
 * which wmpc_osqp_adjoint_f32 equilibrates per instance (symmetric Ruiz
 * scaling, 2 passes), factors once (sigma on P's diagonal, -delta on the
 * equilibrated constraint diagonal) and refines `refine` times against the
 * exact matrix.  Where the rows with a positive multiplier are feasible (s <= 0, a
 * converged solution) the result is the reference's limit; at an iterate
 * that violates such a row the slack is taken as 0, where the reference's U
 * is nearly singular.  Outputs as the reference's mapping of sol
 * (solver.py:1141-1192).  `kkt` is the QDLDL plan of the forward KKT pattern
 * (FixedOSQPPlan.qdldl_plan), referenced, not copied.  Host pointers, read
 * during creation.
 */
typedef struct wmpc_adjoint_kkt_desc {
  int32_t n, m, nnz_p, nnz_a, nnz_kkt;
  int32_t n_l, n_u, n_eq;
  double delta;                   /* regularisation of the equilibrated constraint diagonal */
  const int32_t* kkt_rows;        /* [nnz_kkt] original upper-CSC row of each KKT slot */
  const int32_t* kkt_cols;        /* [nnz_kkt] its column */
  const int32_t* p_to_kkt;        /* [nnz_p] */
  const double* p_diag_sigma;     /* [nnz_p] sigma on diagonal entries, else 0 */
  const int32_t* a_to_kkt;        /* [nnz_a] */
  const int32_t* lower_indices;   /* [n_l] */
  const int32_t* upper_indices;   /* [n_u] */
  const int32_t* eq_indices;      /* [n_eq] */
  const int32_t* p_upper_rows;    /* [nnz_p] */
  const int32_t* p_upper_cols;    /* [nnz_p] */
  const int32_t* a_rows;          /* [nnz_a] */
  const int32_t* a_cols;          /* [nnz_a] */
} wmpc_adjoint_kkt_desc;
int wmpc_adjoint_kkt_plan_create(const wmpc_adjoint_kkt_desc* desc, const wmpc_qdldl_plan* kkt,
                                 wmpc_adjoint_plan** out);
/* 1 for a plan from wmpc_adjoint_kkt_plan_create, 0 otherwise, -1 on NULL. */
int32_t wmpc_adjoint_is_kkt_form(const wmpc_adjoint_plan* plan);
/* Adjoint dimension 2*half (the width of `sol`). */
int32_t wmpc_adjoint_dim(const wmpc_adjoint_plan* plan);
/* Floats of scratch wmpc_osqp_adjoint_f32 needs at instance stride ld:
 * (nnz_l + 4*dim + 2*(n_l + n_u)) * ld, or for the KKT form
 * (nnz_l of the KKT + 5*(n + m) + 3*m) * ld.  -1 on bad input. */
int64_t wmpc_adjoint_work_floats(const wmpc_adjoint_plan* plan, int64_t ld);

/*
 * adjoint_derivative_compute + get_vec + get_mat (solver.py:1047-1192) in one
 * launch, one lane per instance: multipliers and slacks from (x, y), the
 * eps-regularised adjoint matrix assembled and factored, the solve, `refine`
 * refinement steps against the unperturbed matrix (for a KKT-form plan: the
 * system described at wmpc_adjoint_kkt_plan_create), then
 *   dq = rx [n][ld], dl = ryl, du = -ryu [m][ld],
 *   dp [nnz_p][ld] (may be NULL), da [nnz_a][ld] (may be NULL),
 *   sol [dim][ld] (may be NULL), the refined adjoint solution.
 * p_values/a_values: [nnz][pm_ld], pm_ld = 0 when shared by the batch.
 * l, u, y, dy [m][ld]; x, dx [n][ld].  dy is the cotangent of the split
 * multipliers (y_l, y_u), i.e. of |y|, as in the reference.
 * work: work_len >= wmpc_adjoint_work_floats(plan, ld) floats.
 */
int wmpc_osqp_adjoint_f32(const wmpc_adjoint_plan* plan, const float* p_values,
                          const float* a_values, int64_t pm_ld, const float* l, const float* u,
                          const float* x, const float* y, const float* dx, const float* dy,
                          int32_t refine, float* work, int64_t work_len, float* sol, float* dq,
                          float* dl, float* du, float* dp, float* da, int64_t batch, int64_t ld,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* WARPMPC_HIP_H */
