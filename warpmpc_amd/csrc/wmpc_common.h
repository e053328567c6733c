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

inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + " launch: " + hipGetErrorString(e));
    return WMPC_ERR_HIP;
  }
  return WMPC_OK;
}

// Instances per wavefront of the batched kernels for a batch (8, 16, 32 or
// 64): the override set by wmpc_set_instances_per_wave, else 64 once the batch
// gives every CU of the chip a wavefront (kFillWaves wavefronts), else the
// smallest power of two that still does.  The kernels are bound by the scalar
// issue of a CU (shared by its four SIMDs), so more wavefronts per CU than one
// only add scalar work per instance; more CUs add scalar units.
constexpr int64_t kFillWaves = 256;
int32_t instances_per_wave(int64_t batch);
struct QdldlProg;
// ... capped by the LDS columns of a pattern's programs
int32_t instances_per_wave(int64_t batch, const QdldlProg& P);
inline int64_t wave_grid(int64_t batch, int32_t ipw) { return (batch + ipw - 1) / ipw; }

// Largest dynamic LDS a workgroup may request on gfx950 (160 KiB per CU).
constexpr size_t kMaxLdsBytes = 160 * 1024;
// Default LDS budget of one 64-lane workgroup: four workgroups (one wavefront
// per SIMD) share a CU.
constexpr size_t kDefaultLdsBytes = 40 * 1024;

}  // namespace wmpc
