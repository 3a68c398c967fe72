// Calibration kernels of known work (cm_probe_alu, cm_probe_stream): the
// rates the device sustains for plain vector-ALU loops and for streaming
// copies, reads and fills of HBM, so that a kernel's roofline fraction can
// be read against what the chip delivers as well as against its published
// peaks (157.3 TFLOP/s FP32 with packed FMA, 78.6 FP64, 8 TB/s).
//
// ALU loops: every thread runs `iters` trips of 16 independent operations
// (x = x + c, or x = fma(x, a, c)) with a and c kernel arguments, so nothing
// folds at compile time; 16 chains per thread and 16 waves per CU hide the
// ALU latency.  Packed FP32 instructions are disabled for all device code, so
// the FP32 rates are those of one operation per lane per instruction.
// Streams: 16-byte nontemporal vector accesses over a grid-stride loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "../../include/compmatmul_amd.h"
#include "runtime.hpp"

namespace cm {

// Results the probes never store in practice; as device globals visible to
// the host they keep the compiler from removing the loops.
__device__ float g_probe_sink_f;
__device__ double g_probe_sink_d;

namespace {

template <typename T>
__device__ __forceinline__ void sink(T v);
template <>
__device__ __forceinline__ void sink<float>(float v) { g_probe_sink_f = v; }
template <>
__device__ __forceinline__ void sink<double>(double v) { g_probe_sink_d = v; }

template <typename T, bool FMA>
__global__ __launch_bounds__(256) void k_alu_loop(int iters, T a, T c) {
  T x[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) x[j] = T(threadIdx.x + j);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = FMA ? fma(x[j], a, c) : x[j] + c;
  }
  T s = 0;
#pragma unroll
  for (int j = 0; j < 16; ++j) s += x[j];
  // the inputs keep every x[j] positive, so this store never runs; it keeps
  // the loop's results live
  if (s < T(0)) sink<T>(s);
}

constexpr int kStreamThreads = 256;
typedef float v4f __attribute__((ext_vector_type(4)));

// Grid-stride streams of 16-byte nontemporal accesses, U per thread per
// trip.  Launch shapes measured best on MI355X over 1 GiB (unroll 1, 2, 4,
// 8, 16; 512 .. 16384 workgroups; plain or nontemporal): copy U = 2 with 3
// workgroups per CU (about 6.0 TB/s read + write), read U = 1 with 6 per CU
// (about 7.4 TB/s), fill U = 1 with 3 per CU (about 6.3 TB/s); the
// runtime's own device-to-device copy reached about 5.1 TB/s.
template <int U>
__global__ __launch_bounds__(kStreamThreads) void k_stream_copy(const v4f* __restrict__ src,
                                                                v4f* __restrict__ dst,
                                                                size_t n4) {
  const size_t stride = size_t(gridDim.x) * kStreamThreads;
  size_t i = size_t(blockIdx.x) * kStreamThreads + threadIdx.x;
  for (; i + (U - 1) * stride < n4; i += U * stride) {
    v4f v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < U; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
  }
  for (; i < n4; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ __launch_bounds__(kStreamThreads) void k_stream_read(const v4f* __restrict__ src,
                                                                size_t n4) {
  const size_t stride = size_t(gridDim.x) * kStreamThreads;
  float acc = 0.f;
  for (size_t i = size_t(blockIdx.x) * kStreamThreads + threadIdx.x; i < n4; i += stride) {
    const v4f v = __builtin_nontemporal_load(src + i);
    acc += v.x + v.y + v.z + v.w;
  }
  // never true for finite data; keeps the loads live
  if (acc != acc) g_probe_sink_f = acc;
}

__global__ __launch_bounds__(kStreamThreads) void k_stream_fill(v4f* __restrict__ dst,
                                                                size_t n4, float value) {
  const size_t stride = size_t(gridDim.x) * kStreamThreads;
  const v4f v = {value, value, value, value};
  for (size_t i = size_t(blockIdx.x) * kStreamThreads + threadIdx.x; i < n4; i += stride)
    __builtin_nontemporal_store(v, dst + i);
}

// true when [p, p + bytes) lies inside one device allocation
bool device_range(const void* p, size_t bytes) {
  if (!p || (reinterpret_cast<uintptr_t>(p) & 15u)) return false;
  hipPointerAttribute_t at{};
  if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type != hipMemoryTypeDevice) {
    (void)hipGetLastError();
    return false;
  }
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  const uintptr_t lo = reinterpret_cast<uintptr_t>(base), q = reinterpret_cast<uintptr_t>(p);
  return q >= lo && bytes <= size && q - lo <= size - bytes;
}

// one untimed launch, then `reps` launches between two events
template <typename F>
int time_launches(int reps, F&& launch, double* ms) {
  hipStream_t s = rt::stream();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return rt::error(CM_EDEVICE, "probe: hipEventCreate failed");
  }
  hipError_t err = launch(s);
  if (err == hipSuccess) err = hipEventRecord(e0, s);
  for (int r = 0; r < reps && err == hipSuccess; ++r) err = launch(s);
  if (err == hipSuccess) err = hipEventRecord(e1, s);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float t = 0.f;
  if (err == hipSuccess) err = hipEventElapsedTime(&t, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (err != hipSuccess) return rt::error(CM_EDEVICE, std::string("probe: ") + hipGetErrorString(err));
  *ms = double(t) / reps;
  return CM_OK;
}

}  // namespace

}  // namespace cm

extern "C" int cm_probe_alu(int kind, int blocks, int iters, int reps, double* ms, double* flops) {
  using namespace cm;
  std::lock_guard<std::recursive_mutex> lk(rt::mutex());
  // at most 2^26 workgroup trips per launch (about 10 ms at the FP64 FMA
  // rate) and 100 timed launches, so no call runs for long
  if (kind < CM_PROBE_F32_ADD || kind > CM_PROBE_F64_FMA || blocks < 1 || iters < 1 ||
      int64_t(blocks) * iters > (int64_t(1) << 26) || reps < 1 || reps > 100 || !ms || !flops)
    return rt::error(CM_EINVAL, "cm_probe_alu: invalid argument");
  int rc = rt::init();
  if (rc) return rc;
  auto launch = [&](hipStream_t s) -> hipError_t {
    const dim3 g(blocks), t(256);
    switch (kind) {
      case CM_PROBE_F32_ADD: k_alu_loop<float, false><<<g, t, 0, s>>>(iters, 1.0f, 1e-7f); break;
      case CM_PROBE_F32_FMA: k_alu_loop<float, true><<<g, t, 0, s>>>(iters, 1.0f, 1e-7f); break;
      case CM_PROBE_F64_ADD: k_alu_loop<double, false><<<g, t, 0, s>>>(iters, 1.0, 1e-15); break;
      default: k_alu_loop<double, true><<<g, t, 0, s>>>(iters, 1.0, 1e-15); break;
    }
    return hipGetLastError();
  };
  rc = time_launches(reps, launch, ms);
  if (rc) return rc;
  const bool fma = kind == CM_PROBE_F32_FMA || kind == CM_PROBE_F64_FMA;
  *flops = double(blocks) * 256.0 * double(iters) * 16.0 * (fma ? 2.0 : 1.0);
  return CM_OK;
}

extern "C" int cm_probe_stream(int kind, void* dst, const void* src, size_t bytes, int reps,
                               double* ms, double* moved) {
  using namespace cm;
  std::lock_guard<std::recursive_mutex> lk(rt::mutex());
  if (kind < CM_PROBE_COPY || kind > CM_PROBE_FILL || bytes == 0 || bytes % 16 || reps < 1 ||
      reps > 100 || !ms || !moved)
    return rt::error(CM_EINVAL, "cm_probe_stream: invalid argument");
  int rc = rt::init();
  if (rc) return rc;
  const bool reads = kind != CM_PROBE_FILL, writes = kind != CM_PROBE_READ;
  if ((reads && !device_range(src, bytes)) || (writes && !device_range(dst, bytes)))
    return rt::error(CM_EINVAL,
                     "cm_probe_stream: buffers must be 16-byte aligned device memory of `bytes`");
  if (kind == CM_PROBE_COPY) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(dst), b = reinterpret_cast<uintptr_t>(src);
    if ((a > b ? a - b : b - a) < bytes)
      return rt::error(CM_EINVAL, "cm_probe_stream: copy buffers overlap");
  }
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return rt::error(CM_EDEVICE, "cm_probe_stream: device query failed");
  const size_t n4 = bytes / 16;
  const size_t per_cu = kind == CM_PROBE_READ ? 6 : 3;
  const unsigned grid = unsigned(std::max<size_t>(
      1, std::min<size_t>(size_t(cus) * per_cu, (n4 + kStreamThreads - 1) / kStreamThreads)));
  auto launch = [&](hipStream_t s) -> hipError_t {
    switch (kind) {
      case CM_PROBE_COPY:
        k_stream_copy<2><<<grid, kStreamThreads, 0, s>>>(static_cast<const v4f*>(src),
                                                         static_cast<v4f*>(dst), n4);
        break;
      case CM_PROBE_READ:
        k_stream_read<<<grid, kStreamThreads, 0, s>>>(static_cast<const v4f*>(src), n4);
        break;
      default:
        k_stream_fill<<<grid, kStreamThreads, 0, s>>>(static_cast<v4f*>(dst), n4, 0.5f);
        break;
    }
    return hipGetLastError();
  };
  rc = time_launches(reps, launch, ms);
  if (rc) return rc;
  *moved = double(bytes) * (kind == CM_PROBE_COPY ? 2.0 : 1.0);
  return CM_OK;
}
