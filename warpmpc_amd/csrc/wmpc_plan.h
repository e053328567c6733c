// Host-side definition of the opaque QDLDL plan handle shared by the library's
// translation units.
#pragma once

#include <vector>

#include "wmpc_program.h"
#include "wmpc_qdldl_device.h"

struct wmpc_qdldl_raw;  // tables of the table-direct kernels (wmpc_qdldl_raw.hip)

struct wmpc_qdldl_plan {
  wmpc::QdldlProg prog;    // device programs
  wmpc::HostProgram host;  // host copy of the same programs
  // copy of the ragged pattern (re-derivation for another LDS budget)
  std::vector<std::vector<int32_t>> arrays;
  wmpc_qdldl_pattern pattern{};
  void* block = nullptr;
  int device = 0;
  size_t lds_budget = wmpc::kDefaultLdsBytes;
  int32_t ipw = 0;  // per-plan instances per wavefront (0: process-wide choice)
  // incremented whenever the programs are re-derived (and their device tables
  // replaced): holders of a copy of `prog` compare it to detect a stale copy
  uint64_t generation = 0;
  wmpc_qdldl_raw* raw = nullptr;  // float64 and level-scheduled kernels' tables
};

namespace wmpc {
int raw_upload(const wmpc_qdldl_pattern& pt, wmpc_qdldl_raw** out);
void raw_destroy(wmpc_qdldl_raw* raw);
}  // namespace wmpc
