// Internal interface of the batch engine (batch.cpp) for the request server
// (serve.cpp): plans into caller-owned device memory on a caller's stream.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/compmatmul_amd.h"

namespace cm {

// Storage of a product's hash arrays (d x m, d x p) and sketch (d x b) kept
// outside the plan's arena, so it can outlive the plan.
struct Retained {
  uint32_t* hA;
  uint32_t* hB;
  float* P;
};

struct PlanEnv {
  hipStream_t stream = nullptr;      // uploads and runs; nullptr: the library stream
  char* arena = nullptr;             // caller-owned arena; nullptr: allocate one
  size_t arena_bytes = 0;
  const Retained* keep = nullptr;    // per product, or nullptr: carved from the arena
  size_t* size_only = nullptr;       // set: only compute the arena size, plan nothing
  // pinned host memory for the table and hash-parameter uploads, reused by
  // the caller only once this plan's stream has passed them: uploads that
  // fit are issued from it without waiting; nullptr or too small: uploaded
  // from pageable memory and waited for
  char* staging = nullptr;
  size_t staging_bytes = 0;
};

// What a finished run leaves for product i (device pointers).
struct ProductView {
  const unsigned long long* counter;  // recovered entries found
  const int32_t* oi;
  const int32_t* oj;
  const float* ov;
};

int batch_plan_env(int count, const cm_product* ps, unsigned flags, int grain, const PlanEnv& env,
                   cm_batch** out);
int batch_enqueue(cm_batch* bt);  // on the plan's stream; timed there when timing is on
ProductView batch_product(const cm_batch* bt, int i);
// the recovery counters of products 0 .. count-1, contiguous (device)
const unsigned long long* batch_counters(const cm_batch* bt);
void batch_destroy(cm_batch* bt);  // host state; the arena too when the plan allocated it

}  // namespace cm
