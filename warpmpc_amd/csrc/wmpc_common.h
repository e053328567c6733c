// Shared host/device definitions of the warpmpc HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/warpmpc_hip.h"

namespace wmpc {

// Per-thread last-error text behind wmpc_last_error().
void set_error(const std::string& msg);
void clear_error();

#define WMPC_HIP_CHECK(expr)                                                     \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ::wmpc::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));     \
      return WMPC_ERR_HIP;                                                       \
    }                                                                            \
  } while (0)

#define WMPC_REQUIRE(cond, msg)             \
  do {                                      \
    if (!(cond)) {                          \
      ::wmpc::set_error(msg);               \
      return WMPC_ERR_INVALID;              \
    }                                       \
  } while (0)

// Device-resident program tables of one QDLDL pattern.  All int32, one
// allocation (`block`).  Row k of the factorization:
//   reach slots s = 0..R-1 correspond to proc entries r = proc_ptr[k] + s;
//   init entry j adds a[init_src[j]] into slot init_slot[j];
//   proc entry r (column c = proc_col[r]) first applies the updates
//   y[upd_slot[u]] -= L[lp[c] + (u - upd_ptr[r])] * y[s] for u in
//   [upd_ptr[r], upd_ptr[r+1]), then writes L[k, c] to slot proc_lidx[r].
struct QdldlTables {
  int32_t n = 0, nnz_a = 0, nnz_l = 0, max_reach = 0, n_proc = 0, n_upd = 0;
  const int32_t* diag_src = nullptr;  // [n] original value slot of A_kk
  const int32_t* init_ptr = nullptr;  // [n+1]
  const int32_t* init_slot = nullptr;
  const int32_t* init_src = nullptr;  // original value slot
  const int32_t* proc_ptr = nullptr;  // [n+1]
  const int32_t* proc_col = nullptr;
  const int32_t* proc_lidx = nullptr;
  const int32_t* upd_ptr = nullptr;   // [n_proc+1]
  const int32_t* upd_slot = nullptr;
  const int32_t* lp = nullptr;        // [n+1]
  const int32_t* li = nullptr;        // [nnz_l]
  const int32_t* p = nullptr;         // [n]
  const int32_t* pinv = nullptr;      // [n]
};

}  // namespace wmpc

struct wmpc_qdldl_plan {
  wmpc::QdldlTables t;
  void* block = nullptr;
  int device = 0;
};
