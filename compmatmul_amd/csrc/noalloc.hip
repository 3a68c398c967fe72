// Fixed device scratch of the no-allocation entry point (cm_matmul_noalloc).
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc) get every
// intermediate of a product from this one array, CM_NOALLOC_SCRATCH_BYTES
// per GPU, held for the life of the process.  In host code `scratch` names
// a host-side stand-in, not device memory: its device address is looked up
// once per device with hipGetSymbolAddress (with that device current) and
// only that address is ever passed to kernels.  Every call on a GPU shares
// the array and finds the previous call's contents in it; nothing in it is
// assumed to start at zero.
#include <hip/hip_runtime.h>

#include "../../include/compmatmul_amd.h"
#include "kernels.hpp"

namespace cm {

namespace {

constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[CM_NOALLOC_SCRATCH_BYTES];

void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device

}  // namespace

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return -1;
  if (g_scratch[device]) return 0;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return -1;
  g_scratch[device] = ptr;
  return 0;
}

int noalloc_scratch(int device, void** ptr, size_t* bytes) {
  if (device < 0 || device >= kMaxDevices || !g_scratch[device]) return -1;
  *ptr = g_scratch[device];
  *bytes = CM_NOALLOC_SCRATCH_BYTES;
  return 0;
}

}  // namespace cm
