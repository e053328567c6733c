/*
 * CPU oracle of the batched fixed-pattern QDLDL factor/solve and OSQP ADMM.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the tests, the smoke test
 * and the CPU-baseline leg of the benchmark compare the HIP library against; no
 * product code path links or calls it.
 *
 * It restates, per instance and in the same operation order:
 *   - the up-looking numeric factorization of warpmpc/jax_qdldl/core.py:867-933
 *     (== warp_backend.py:41-101): row k gathers column k of P A Pᵀ into y,
 *     walks the etree reach in QDLDL order, L[k,c] = y_c * Dinv_c,
 *     D_k = A_kk - sum y_c L[k,c], Dinv_k = 1 / D_k;
 *   - the column-oriented triangular solves of core.py:935-1028
 *     (== warp_backend.py:176-218): x = rhs[p], forward column scatter,
 *     x *= Dinv, backward column gather, out = x[pinv];
 *   - the fixed-iteration OSQP ADMM of warpmpc/jax_osqp/solver.py:764-872:
 *     KKT value scatter, input scaling, max_iter iterations of _iterate,
 *     unscaling, residuals and objective;
 *   - OSQP's adjoint derivative of warpmpc/jax_osqp/solver.py:1035-1192
 *     (adjoint_oracle_impl.h).
 *
 * Compiled twice from qdldl_oracle_impl.h: suffix _f64 (REAL = double, the
 * oracle proper) and _f32 (REAL = float, the FP32 CPU counterpart of the
 * reference's qdldl_backend="jax" path used as the benchmark's CPU baseline).
 * Built with -ffp-contract=off so that products and sums round separately, as
 * XLA's scatter-add formulation does.  Instances run under OpenMP.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REAL double
#define SFX(name) name##_f64
#include "qdldl_oracle_impl.h"
#include "adjoint_oracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX(name) name##_f32
#include "qdldl_oracle_impl.h"
#include "adjoint_oracle_impl.h"
#undef REAL
#undef SFX

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Threads of the following parallel regions (n >= 1). */
void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  if (n >= 1) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
