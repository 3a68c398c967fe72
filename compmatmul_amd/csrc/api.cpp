// C-ABI of the MI355X compressed matrix product (declared in
// include/compmatmul_amd.h).  Host orchestration: argument checks, hash
// parameter drawing, workspace management, kernel sequencing on one HIP
// stream, timing.  All device arithmetic lives in the .hip files.
#include "../../include/compmatmul_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <initializer_list>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "host_copy.hpp"
#include "host_hash.hpp"
#include "recover_sort.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

struct cm_sketch {
  int fwht, f64, d, b, lb, m, p, family;
  uint64_t seed;
  uint32_t* hA;  // device, d x m packed bucket/sign
  uint32_t* hB;  // device, d x p
  void* P;       // device, d x b sketch values (element type per f64)
  bool on_default = true;  // made on the library's default context (its stream)
  int device = 0;          // where hA, hB and P live
};

namespace {

thread_local std::string g_err;
std::recursive_mutex g_mu;

// smallest log2(b) sent to the register kernels (padding to 2048 costs at
// most 2x there; smaller sketches use the LDS kernel)
constexpr int kRegMinLb = 10;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define CM_HIP(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return fail(e_ == hipErrorOutOfMemory ? CM_ENOMEM : CM_EDEVICE,             \
                  std::string(#expr) + ": " + hipGetErrorString(e_));             \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
};

struct State {
  int device = -1;
  hipStream_t stream = nullptr;
  DevBuf partial, V, params, tables, AT, B, C, counter, oi, oj, ov, plainA, plainB, entries, scale;
  void* fft_t2 = nullptr;  // W_2048^(L bitrev5(q)) table of the FFT register kernel
  void* pinned = nullptr;
  size_t pinned_n = 0;
  void* xfer = nullptr;  // two pinned transfer slots (kXferChunk each)
  hipEvent_t xfer_ev[2] = {nullptr, nullptr};
  std::map<std::pair<int, int>, void*> twiddles;  // (log2 b, f64) -> table
  bool timing = false;
  std::string timing_only;  // empty: every kernel class
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<std::string, std::pair<double, int64_t>> totals;
  // a context's workspace: every per-call device intermediate is carved
  // from [ws_base, ws_base + ws_cap) (nullptr: buffers grow on demand)
  char* ws_base = nullptr;
  size_t ws_cap = 0;
};
// The library's default context (device 0 on first use), which every entry
// point without a context argument uses, and the context of this thread's
// current call.
State g_default;
thread_local State* g_cur = &g_default;
inline State& cur() { return *g_cur; }

int ensure_buf(DevBuf& b, size_t bytes) {
  if (b.n >= bytes && b.p) return CM_OK;
  if (b.p) {
    CM_HIP(hipStreamSynchronize(cur().stream));
    CM_HIP(hipFree(b.p));
    b.p = nullptr;
    b.n = 0;
  }
  const size_t want = std::max<size_t>(bytes, 256);
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(CM_ENOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
  }
  b.n = want;
  return CM_OK;
}

int ensure_pinned(size_t bytes) {
  if (cur().pinned_n >= bytes) return CM_OK;
  if (cur().pinned) {
    CM_HIP(hipStreamSynchronize(cur().stream));
    CM_HIP(hipHostFree(cur().pinned));
    cur().pinned = nullptr;
    cur().pinned_n = 0;
  }
  CM_HIP(hipHostMalloc(&cur().pinned, bytes, hipHostMallocDefault));
  cur().pinned_n = bytes;
  return CM_OK;
}

// Large transfers between pageable host memory and the device go through
// two pinned slots, the host copy of one chunk overlapping the DMA of the
// other: pageable copies run at about 31 GB/s to the device and 18 GB/s
// back on MI355X, pinned DMA at about 56 GB/s; the host copies run on a few
// threads (host_copy.hpp).  The caller's host memory is read or written by
// the host only, so it may be reused on return.  Chunks of 1 MiB up to
// 4 MiB transfers (the overlap needs several chunks), 4 MiB above (measured
// with the threaded copies: 17 MB operands 8% faster than with 2 MiB, no
// gain at 8 MiB).
constexpr size_t kXferChunk = size_t(4) << 20;  // slot size, the largest chunk
constexpr size_t kXferMin = size_t(1) << 20;    // below this, one pageable copy
inline size_t xfer_chunk(size_t bytes) { return bytes <= (size_t(4) << 20) ? size_t(1) << 20 : kXferChunk; }

int ensure_xfer() {
  State& st = cur();
  if (st.xfer && st.xfer_ev[1]) return CM_OK;
  if (!st.xfer) CM_HIP(hipHostMalloc(&st.xfer, 2 * kXferChunk, hipHostMallocDefault));
  for (hipEvent_t& e : st.xfer_ev)
    if (!e) CM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return CM_OK;
}

int h2d(void* dev, const void* host, size_t bytes) {
  if (bytes < kXferMin) {
    CM_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, cur().stream));
    return CM_OK;
  }
  if (int rc = ensure_xfer()) return rc;
  State& st = cur();
  const size_t chunk = xfer_chunk(bytes);
  for (size_t off = 0, i = 0; off < bytes; off += chunk, ++i) {
    const size_t len = std::min(chunk, bytes - off);
    char* slot = static_cast<char*>(st.xfer) + (i & 1) * kXferChunk;
    CM_HIP(hipEventSynchronize(st.xfer_ev[i & 1]));  // the slot's last DMA is done
    cm::parallel_copy(slot, static_cast<const char*>(host) + off, len);
    CM_HIP(hipMemcpyAsync(static_cast<char*>(dev) + off, slot, len, hipMemcpyHostToDevice, st.stream));
    CM_HIP(hipEventRecord(st.xfer_ev[i & 1], st.stream));
  }
  return CM_OK;
}

// device to host, complete on return
int d2h(void* host, const void* dev, size_t bytes) {
  State& st = cur();
  if (bytes < kXferMin) {
    CM_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, st.stream));
    CM_HIP(hipStreamSynchronize(st.stream));
    return CM_OK;
  }
  if (int rc = ensure_xfer()) return rc;
  const size_t chunk = xfer_chunk(bytes);
  const size_t n = (bytes + chunk - 1) / chunk;
  auto issue = [&](size_t i) -> hipError_t {
    const size_t off = i * chunk, len = std::min(chunk, bytes - off);
    char* slot = static_cast<char*>(st.xfer) + (i & 1) * kXferChunk;
    hipError_t e = hipMemcpyAsync(slot, static_cast<const char*>(dev) + off, len, hipMemcpyDeviceToHost, st.stream);
    return e == hipSuccess ? hipEventRecord(st.xfer_ev[i & 1], st.stream) : e;
  };
  CM_HIP(issue(0));
  for (size_t i = 0; i < n; ++i) {
    if (i + 1 < n) CM_HIP(issue(i + 1));  // its slot's chunk i - 1 was copied out
    CM_HIP(hipEventSynchronize(st.xfer_ev[i & 1]));
    const size_t off = i * chunk, len = std::min(chunk, bytes - off);
    cm::parallel_copy(static_cast<char*>(host) + off, static_cast<const char*>(st.xfer) + (i & 1) * kXferChunk, len);
  }
  return CM_OK;
}

int init_locked(int device) {
  if (cur().stream && (device < 0 || device == cur().device)) return CM_OK;
  if (device < 0) device = 0;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(CM_EDEVICE, "no HIP device is available (compmatmul_amd needs an MI355X)");
  if (device >= count) return fail(CM_EINVAL, "device index out of range");
  hipDeviceProp_t prop;
  CM_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(CM_EDEVICE, std::string("compmatmul_amd is built for gfx950, found ") +
                                prop.gcnArchName);
  CM_HIP(hipSetDevice(device));
  if (cur().stream) CM_HIP(hipStreamDestroy(cur().stream));
  CM_HIP(hipStreamCreateWithFlags(&cur().stream, hipStreamNonBlocking));
  cur().device = device;
  // the no-allocation entry's scratch address on this device
  const int rc = cm::noalloc_setup(device);
  if (rc) return fail(CM_EDEVICE, "no-allocation scratch: hipGetSymbolAddress failed");
  return CM_OK;
}

hipEvent_t get_event() {
  if (!cur().event_pool.empty()) {
    hipEvent_t e = cur().event_pool.back();
    cur().event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;  // that launch goes untimed
  return e;
}

// launch wrapper: brackets the launch with events on `stream` (the stream
// the launch runs on) when timing is on
template <class F>
int timed_on(hipStream_t stream, const char* name, F&& launch) {
  hipEvent_t a = nullptr, b = nullptr;
  const std::string& only = cur().timing_only;
  const bool prefix = !only.empty() && only.back() == '*';  // "compress*": every compress class
  bool on = cur().timing && (only.empty() || only == name ||
                         (prefix && std::strncmp(name, only.c_str(), only.size() - 1) == 0));
  if (on) {
    a = get_event();
    b = get_event();
    on = a && b && hipEventRecord(a, stream) == hipSuccess;
    if (!on) {
      if (a) cur().event_pool.push_back(a);
      if (b) cur().event_pool.push_back(b);
    }
  }
  const hipError_t e = launch();
  if (on) {
    if (hipEventRecord(b, stream) == hipSuccess) {
      cur().pending.push_back({name, a, b});
    } else {
      cur().event_pool.push_back(a);
      cur().event_pool.push_back(b);
    }
  }
  if (e != hipSuccess) return fail(CM_EDEVICE, std::string(name) + " launch failed: " + hipGetErrorString(e));
  return CM_OK;
}

template <class F>
int timed(const char* name, F&& launch) {
  return timed_on(cur().stream, name, launch);
}

// each pair's own end event is waited for: pairs may sit on other streams
void resolve_timing() {
  if (cur().pending.empty()) return;
  for (auto& r : cur().pending) {
    float ms = 0.f;
    (void)hipEventSynchronize(r.b);
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    auto& t = cur().totals[r.name];
    t.first += ms;
    t.second += 1;
    cur().event_pool.push_back(r.a);
    cur().event_pool.push_back(r.b);
  }
  cur().pending.clear();
}

int ilog2(int64_t x) {
  int l = 0;
  while ((int64_t(1) << (l + 1)) <= x) ++l;
  return l;
}

int check_db(int d, int b) {
  if (b < 2 || (b & (b - 1)) != 0)
    return fail(CM_EINVAL, "b must be power of 2 greater than 1 (got " + std::to_string(b) + ")");
  if (d < 1) return fail(CM_EINVAL, "d must be positive (got " + std::to_string(d) + ")");
  return CM_OK;
}

int check_family(int family) {
  if (family < 0 || family > 3) return fail(CM_EHASH, "Tried to use an invalid hash");
  return CM_OK;
}

// W_b^j = exp(-2 pi i j / b) for j < b (tables.hip), built once per (b,
// element type) and kept for the process
int twiddles(int lb, int f64, const void** out) {
  auto key = std::make_pair(lb, f64);
  auto it = cur().twiddles.find(key);
  if (it != cur().twiddles.end()) {
    *out = it->second;
    return CM_OK;
  }
  void* dev = nullptr;
  CM_HIP(hipMalloc(&dev, cm::twiddle_bytes(lb, f64)));
  hipError_t e = cm::launch_twiddles(lb, f64, dev, cur().stream);
  // batches and server waves read the table from streams of their own
  if (e == hipSuccess) e = hipStreamSynchronize(cur().stream);
  if (e != hipSuccess) {
    (void)hipFree(dev);
    return fail(CM_EDEVICE, std::string("twiddle table: ") + hipGetErrorString(e));
  }
  cur().twiddles[key] = dev;
  *out = dev;
  return CM_OK;
}

// Table of the FFT register kernel: W_2048^(L * bitrev5(q)) at [q*64 + L].
int fft_t2_table(const void** out) {
  if (!cur().fft_t2) {
    void* dev = nullptr;
    CM_HIP(hipMalloc(&dev, 2048 * sizeof(float2)));
    hipError_t e = cm::launch_t2_table(dev, cur().stream);
    if (e == hipSuccess) e = hipStreamSynchronize(cur().stream);
    if (e != hipSuccess) {
      (void)hipFree(dev);
      return fail(CM_EDEVICE, std::string("DFT table: ") + hipGetErrorString(e));
    }
    cur().fft_t2 = dev;
  }
  *out = cur().fft_t2;
  return CM_OK;
}

// Device memory of one product: the library's growing buffers, or, for the
// no-allocation entry, a caller-provided region carved front to back (the
// twiddle tables are then written into it too).
struct Workspace {
  char* base = nullptr;  // nullptr: the library's buffers and cached tables
  size_t cap = 0, off = 0;
  const char* what = "no-allocation scratch";
  bool carved() const { return base != nullptr; }
  static size_t up(size_t x) { return (x + 255) & ~size_t(255); }
  // bytes still free in the region (no limit for the library's buffers)
  size_t room() const { return carved() ? (up(off) < cap ? cap - up(off) : 0) : SIZE_MAX; }
  int take(DevBuf& buf, size_t bytes, void** out) {
    if (!carved()) {
      int rc = ensure_buf(buf, bytes);
      if (rc) return rc;
      *out = buf.p;
      return CM_OK;
    }
    const size_t at = up(off);
    if (at + bytes > cap)
      return fail(CM_ENOMEM, "the product needs more than the " + std::to_string(cap >> 20) + " MiB " +
                                 what + " (at least " +
                                 std::to_string((at + bytes + ((1 << 20) - 1)) >> 20) + " MiB)");
    *out = base + at;
    off = at + bytes;
    return CM_OK;
  }
  int twiddles(int lb, int f64, const void** out) {
    if (!carved()) return ::twiddles(lb, f64, out);
    DevBuf none;
    void* p = nullptr;
    int rc = take(none, cm::twiddle_bytes(lb, f64), &p);
    if (rc) return rc;
    CM_HIP(cm::launch_twiddles(lb, f64, p, cur().stream));
    *out = p;
    return CM_OK;
  }
  int t2_table(const void** out) {
    if (!carved()) return ::fft_t2_table(out);
    DevBuf none;
    void* p = nullptr;
    int rc = take(none, 2048 * sizeof(float2), &p);
    if (rc) return rc;
    CM_HIP(cm::launch_t2_table(p, cur().stream));
    *out = p;
    return CM_OK;
  }
};

// The workspace of one call on the current context: its budget region,
// carved from the start, or the library's buffers.
Workspace call_workspace() {
  Workspace ws;
  if (cur().ws_base) {
    ws.base = cur().ws_base;
    ws.cap = cur().ws_cap;
    ws.what = "context workspace";
  }
  return ws;
}

// Draw h1, h2, s1, s2 and evaluate the packed arrays on the device.
// With no_alloc, only parameters that travel inside the kernel arguments are
// taken (multiply families with 4 d words <= 256), so nothing is allocated
// or uploaded.
// Parameters and tables that do not fit the kernel arguments are uploaded
// into buffers taken from `ws` (a carved workspace: from its region, through
// a synchronous copy instead of the pinned staging buffer).
int build_hashes(int family, int d, int lb, uint64_t seed, int m, int p, uint32_t* outA,
                 uint32_t* outB, Workspace& ws, bool no_alloc = false) {
  cm::HashArgs ha{};
  ha.family = family;
  ha.d = d;
  ha.lb = lb;
  ha.m = m;
  ha.p = p;
  ha.outA = outA;
  ha.outB = outB;
  const bool mul = family == cm::kMultiplyAddShift || family == cm::kMultiplyShift;
  const int64_t words = mul ? int64_t(4) * d * cm::param_words(family) : 0;
  if (mul && words <= 256) {
    ha.use_inline = 1;
    for (int f = 0; f < 4; ++f)
      cm::draw_multiply_into(family, d, seed + uint64_t(f),
                             ha.inline_params + int64_t(f) * d * cm::param_words(family));
  } else if (no_alloc) {
    return fail(CM_ENOTSUP, "the no-allocation entry takes the multiply-add-shift family with "
                            "d <= 32 or multiply-shift with d <= 64");
  } else if (mul) {
    std::vector<uint64_t> all, one;
    for (int f = 0; f < 4; ++f) {
      cm::draw_multiply(family, d, seed + uint64_t(f), one);
      all.insert(all.end(), one.begin(), one.end());
    }
    {
      void* dst = nullptr;
      int rc = ws.take(cur().params, all.size() * 8, &dst);
      if (rc) return rc;
      if (ws.carved()) {
        CM_HIP(hipMemcpyAsync(dst, all.data(), all.size() * 8, hipMemcpyHostToDevice, cur().stream));
        CM_HIP(hipStreamSynchronize(cur().stream));  // `all` is released on return
      } else {
        rc = ensure_pinned(all.size() * 8);
        if (rc) return rc;
        CM_HIP(hipStreamSynchronize(cur().stream));
        std::memcpy(cur().pinned, all.data(), all.size() * 8);
        CM_HIP(hipMemcpyAsync(dst, cur().pinned, all.size() * 8, hipMemcpyHostToDevice, cur().stream));
      }
      ha.mul_params = static_cast<const uint64_t*>(dst);
    }
  } else {
    const int chars = family == cm::kTab8 ? 4 : 2;
    const int64_t full = family == cm::kTab8 ? 256 : 65536;
    const int64_t maxkey = std::max(1, std::max(m, p)) - 1;
    const int64_t tab_len = family == cm::kTab8 ? 256 : std::min<int64_t>(full, maxkey + 1);
    const size_t bytes = size_t(4) * d * chars * tab_len * 4;
    void* dev = nullptr;
    int rc = ws.take(cur().tables, bytes, &dev);
    if (rc) return rc;
    std::vector<uint32_t> staged;  // a carved workspace: host staging, copied synchronously
    if (ws.carved()) {
      staged.resize(bytes / 4);
    } else {
      rc = ensure_pinned(bytes);
      if (rc) return rc;
      CM_HIP(hipStreamSynchronize(cur().stream));
    }
    uint32_t* const host = ws.carved() ? staged.data() : static_cast<uint32_t*>(cur().pinned);
    uint32_t* dst = host;
    std::vector<uint32_t> tabs;
    for (int f = 0; f < 4; ++f) {
      cm::draw_tables(family, d, seed + uint64_t(f), tabs);
      for (int t = 0; t < d; ++t)
        for (int c = 0; c < chars; ++c) {
          const uint32_t* src = tabs.data() + (int64_t(t) * chars + c) * full;
          std::memcpy(dst, src, tab_len * 4);
          dst += tab_len;
        }
    }
    CM_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, cur().stream));
    if (ws.carved()) CM_HIP(hipStreamSynchronize(cur().stream));
    ha.tables = static_cast<const uint32_t*>(dev);
    ha.tab_len = int(tab_len);
  }
  if (no_alloc) {
    CM_HIP(cm::launch_hash(ha, cur().stream));
    return CM_OK;
  }
  return timed("hash", [&] { return cm::launch_hash(ha, cur().stream); });
}

int alloc_sketch(int fwht, int f64, int d, int b, int m, int p, cm_sketch** out) {
  cm_sketch* sk = new cm_sketch{};
  sk->fwht = fwht;
  sk->f64 = f64;
  sk->d = d;
  sk->b = b;
  sk->lb = ilog2(b);
  sk->m = m;
  sk->p = p;
  const size_t esz = f64 ? 8 : 4;
  if (hipGetDevice(&sk->device) != hipSuccess) sk->device = 0;
  hipError_t e1 = hipMalloc(&sk->hA, std::max<size_t>(4, size_t(d) * m * 4));
  hipError_t e2 = hipMalloc(&sk->hB, std::max<size_t>(4, size_t(d) * p * 4));
  hipError_t e3 = hipMalloc(&sk->P, size_t(d) * b * esz);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
    cm_sketch_free(sk);
    return fail(CM_ENOMEM, "device allocation for the sketch failed");
  }
  *out = sk;
  return CM_OK;
}

// stage a host matrix (rows x cols, leading dimension ld) into a device buffer
int stage_in(Workspace& ws, DevBuf& buf, const void* host, int rows, int cols, int64_t ld, size_t esz,
             const void** dev) {
  if (rows == 0 || cols == 0) {
    *dev = nullptr;
    return CM_OK;
  }
  const size_t elems = size_t(rows - 1) * size_t(ld) + size_t(cols);
  void* d = nullptr;
  int rc = ws.take(buf, elems * esz, &d);
  if (rc) return rc;
  if (ws.carved()) {  // a carved workspace takes no pinned slots
    CM_HIP(hipMemcpyAsync(d, host, elems * esz, hipMemcpyHostToDevice, cur().stream));
  } else if ((rc = h2d(d, host, elems * esz))) {
    return rc;
  }
  *dev = d;
  return CM_OK;
}

// Launch the named stages in order: timed for the library's buffers, plain
// launches (no events) for a carved region.
int run_stages(const Workspace& ws,
               std::initializer_list<std::pair<const char*, std::function<hipError_t()>>> stages) {
  for (const auto& st : stages) {
    if (ws.carved()) {
      const hipError_t e = st.second();
      if (e != hipSuccess) return fail(CM_EDEVICE, std::string(st.first) + " launch failed: " + hipGetErrorString(e));
    } else if (int rc = timed(st.first, st.second)) {
      return rc;
    }
  }
  return CM_OK;
}

int compress_device(cm_sketch* sk, const void* AT, int64_t ld_at, const void* B, int64_t ld_b,
                    int n, unsigned flags, Workspace& ws) {
  const int f64 = sk->f64, fwht = sk->fwht, d = sk->d, lb = sk->lb, m = sk->m, p = sk->p;
  const int64_t b = int64_t(1) << lb;
  const size_t esz = f64 ? 8 : 4;
  const size_t cplx = fwht ? 1 : 2;
  const int max_lb2 = cm::compress_lds_max_lb2(fwht, f64);
  const void* tw = nullptr;
  // FFT: per-k power-of-two balance of the pair-packed rows, computed by
  // the first launch of the compress stage (balance.hip)
  cm::PairScaleArgs psa{f64, AT, ld_at, B, ld_b, m, n, p, nullptr};
  auto with_scales = [&]() { return fwht ? hipSuccess : cm::launch_pair_scale(psa, cur().stream); };
  if (!fwht) {
    int rc = ws.twiddles(lb, f64, &tw);
    if (rc) return rc;
    if ((rc = ws.take(cur().scale, std::max<size_t>(1, size_t(n) * 2 * esz), &psa.out))) return rc;
  }
  if (ws.carved() && (flags & (CM_FLAG_PLAIN | CM_FLAG_BF16_MFMA | CM_FLAG_FP16_MFMA)))
    return fail(CM_ENOTSUP, "the plain and matrix-core paths allocate their own buffers");
  if (flags & CM_FLAG_PLAIN) {
    if (lb > cm::plain_max_lb(fwht, f64))
      return fail(CM_ENOTSUP, "b too large for the plain path (2^" +
                                  std::to_string(cm::plain_max_lb(fwht, f64)) + " at most)");
    // Passes over (repetitions, inner indices): the sketches of one pass
    // take at most 4 GiB; a pass holds whole repetitions when they fit, or a
    // range of k of one repetition.
    const size_t row_bytes = size_t(b) * 2 * esz;  // FFT: complex; FWHT: A and B rows
    const int64_t rows_cap = std::max<int64_t>(1, int64_t((size_t(4) << 30) / row_bytes));
    const int nk = n > 0 ? int(std::min<int64_t>(n, rows_cap)) : 0;
    const int nt = nk == n && n > 0 ? int(std::max<int64_t>(1, std::min<int64_t>(d, rows_cap / n))) : 1;
    // MAC chunks of about 32 rows of k: many short workgroups, so the last
    // round of them is a small share of the stage
    const int kchunks = std::max(1, (nk + 31) / 32);
    const int lb2 = std::min(lb, max_lb2);
    int rc;
    if ((rc = ensure_buf(cur().plainA, std::max<size_t>(1, size_t(nt) * nk * b * cplx * esz))) ||
        (rc = ensure_buf(cur().plainB, fwht ? std::max<size_t>(1, size_t(nt) * nk * b * esz) : 1)) ||
        (rc = ensure_buf(cur().entries, std::max<size_t>(1, size_t(kchunks) * nt * b * cplx * esz))) ||
        (rc = ensure_buf(cur().partial, size_t(d) * b * cplx * esz)) ||
        (rc = ensure_buf(cur().V, size_t(d) * b * cplx * esz)))
      return rc;
    if (n == 0) {  // no pass: the spectral sums are zero
      const size_t words = size_t(d) * b * cplx * esz / 4;
      for (size_t w0 = 0; w0 < words; w0 += size_t(1) << 30) {
        const int cnt = int(std::min<size_t>(size_t(1) << 30, words - w0));
        unsigned* at = static_cast<unsigned*>(cur().partial.p) + w0;
        if ((rc = timed("plain_mac", [&] { return cm::launch_zero_u32(at, cnt, cur().stream); }))) return rc;
      }
    }
    for (int t0 = 0; n > 0 && t0 < d; t0 += nt) {
      for (int k0 = 0; k0 < n; k0 += nk) {
        cm::PlainArgs pa{};
        pa.fwht = fwht;
        pa.f64 = f64;
        pa.AT = AT;
        pa.ld_at = ld_at;
        pa.B = B;
        pa.ld_b = ld_b;
        pa.hA = sk->hA;
        pa.hB = sk->hB;
        pa.m = m;
        pa.n = n;
        pa.p = p;
        pa.d = d;
        pa.lb = lb;
        pa.t0 = t0;
        pa.nt = std::min(nt, d - t0);
        pa.k0 = k0;
        pa.nk = std::min(nk, n - k0);
        pa.tw = tw;
        pa.scale = psa.out;
        pa.SA = cur().plainA.p;
        pa.SB = cur().plainB.p;
        pa.kchunks = kchunks;
        pa.part = cur().entries.p;
        pa.acc = cur().partial.p;
        pa.lb2 = lb2;
        pa.first = k0 == 0;
        const bool first_pass = t0 == 0 && k0 == 0;
        if ((rc = timed("plain_sketch", [&] {
               // the balance scales once, ahead of the first pass's sketch
               const hipError_t e = first_pass ? with_scales() : hipSuccess;
               return e != hipSuccess ? e : cm::launch_plain_sketch(pa, cur().stream);
             })))
          return rc;
        if ((rc = timed("plain_transform", [&] { return cm::launch_plain_transform(pa, cur().stream); }))) return rc;
        if ((rc = timed("plain_mac", [&] { return cm::launch_plain_mac(pa, cur().stream); }))) return rc;
      }
    }
    const int f64_reduce = (flags & CM_FLAG_F64_REDUCE) || n > 16384;
    cm::ReduceArgs ra{fwht, f64, f64_reduce, cur().partial.p, 1, d, lb, lb2, tw, cur().V.p};
    if ((rc = timed("reduce", [&] { return cm::launch_reduce_inverse(ra, cur().stream); }))) return rc;
    cm::CombineArgs ca{fwht, f64, cur().V.p, d, lb, lb2, tw, sk->P, lb};
    return timed("combine", [&] { return cm::launch_combine(ca, cur().stream); });
  }

  if (flags & (CM_FLAG_BF16_MFMA | CM_FLAG_FP16_MFMA)) {
    if ((flags & CM_FLAG_BF16_MFMA) && (flags & CM_FLAG_FP16_MFMA))
      return fail(CM_EINVAL, "CM_FLAG_BF16_MFMA and CM_FLAG_FP16_MFMA exclude each other");
    if (fwht || f64 || !cm::fft_bf16_supported(lb))
      return fail(CM_ENOTSUP, "the matrix-core modes take the FFT variant, float32 and "
                              "b = 16^k with 2^12 <= b <= 2^28");
    // two row buffers of at most 512 MiB each
    const int rows = int(std::max<int64_t>(1, std::min<int64_t>(std::max(n, 1), (int64_t(1) << 29) / (b * 8))));
    const int nchunks = n > 0 ? (n + rows - 1) / rows : 0;
    const void* twb = nullptr;
    int rc;
    if ((rc = twiddles(lb, 0, &twb))) return rc;
    const size_t row_bytes = size_t(rows) * b * 8;
    const size_t chunk_bytes = size_t(d) * b * 8;
    if ((rc = ensure_buf(cur().plainA, row_bytes)) || (rc = ensure_buf(cur().plainB, row_bytes)) ||
        (rc = ensure_buf(cur().partial, std::max<size_t>(1, size_t(nchunks) * chunk_bytes))) ||
        (rc = ensure_buf(cur().V, chunk_bytes)))
      return rc;
    cm::Bf16FftArgs fa{};
    fa.AT = static_cast<const float*>(AT);
    fa.ld_at = ld_at;
    fa.B = static_cast<const float*>(B);
    fa.ld_b = ld_b;
    fa.hA = sk->hA;
    fa.hB = sk->hB;
    fa.m = m;
    fa.n = n;
    fa.p = p;
    fa.d = d;
    fa.lb = lb;
    fa.rows = rows;
    fa.nchunks = nchunks;
    fa.tw = twb;
    fa.scale = static_cast<const float2*>(psa.out);
    fa.Z0 = static_cast<float2*>(cur().plainA.p);
    fa.Z1 = static_cast<float2*>(cur().plainB.p);
    fa.partial = static_cast<float*>(cur().partial.p);
    fa.fp16 = (flags & CM_FLAG_FP16_MFMA) ? 1 : 0;
    if ((rc = timed("compress", [&] {
           const hipError_t e = with_scales();
           return e != hipSuccess ? e : cm::launch_fft_bf16_compress(fa, cur().stream);
         })))
      return rc;
    const int f64_reduce = (flags & CM_FLAG_F64_REDUCE) || n > 16384;
    cm::ReduceArgs ra{0, 0, f64_reduce, cur().partial.p, nchunks, d, lb, cm::kRegLb2, twb, cur().V.p};
    if ((rc = timed("reduce", [&] { return cm::launch_reduce_inverse(ra, cur().stream); }))) return rc;
    cm::CombineArgs ca{0, 0, cur().V.p, d, lb, cm::kRegLb2, twb, sk->P, lb};
    return timed("combine", [&] { return cm::launch_combine(ca, cur().stream); });
  }

  // fast path: FP32 with b >= 1024 runs the register-resident kernels
  // (2048-frequency slices; smaller b zero-padded to one slice)
  const bool reg = !f64 && !(flags & CM_FLAG_LDS_KERNEL) && lb >= kRegMinLb &&
                   (fwht ? cm::fwht_reg_supported(lb, m, p) : cm::fft_reg_supported(lb, m, p));
  if (reg) {
    const int lbv = std::max(lb, cm::kRegLb2);  // transform length actually run
    const int lbs = cm::reg_lbs(lb);
    // with >= 2 slices, workgroups of 2^logg waves share one scatter
    const bool grp = !(flags & CM_FLAG_SINGLE_WAVE) &&
                     (fwht ? cm::fwht_grp_supported(lb, m, p) : cm::fft_grp_supported(lb, m, p));
    const int logg = grp ? (fwht ? cm::fwht_grp_log_g(lb) : cm::fft_grp_log_g(lb)) : 0;
    const int64_t per_chunk = (int64_t(d) << lbs) >> logg;  // workgroups per chunk of k
    const int64_t slots = grp ? (fwht ? cm::fwht_grp_slots(lb) : cm::fft_grp_slots(lb))
                              : fwht ? cm::fwht_reg_slots() : cm::fft_reg_slots();
    const int64_t rounds = std::max<int64_t>(1, (int64_t(n) * per_chunk + slots * 64 - 1) / (slots * 64));
    int64_t nchunks = std::max<int64_t>(1, slots * rounds / per_chunk);
    nchunks = std::min<int64_t>(nchunks, std::max(1, n));
    const size_t chunk_bytes = (size_t(d) << lbv) * cplx * esz;
    const void* twv = nullptr;  // twiddles of the transform length run
    const void* t2 = nullptr;
    int rc;
    if (!fwht) {
      if (lbv == lb)
        twv = tw;
      else if ((rc = ws.twiddles(lbv, 0, &twv)))
        return rc;
      if ((rc = ws.t2_table(&t2))) return rc;
    }
    const size_t lists_bytes = grp ? (fwht ? cm::fwht_grp_lists_bytes(d, lb, m, p)
                                           : cm::fft_grp_lists_bytes(d, lb, m, p))
                               : fwht ? cm::fwht_reg_lists_bytes(d, lb, m, p)
                                      : cm::fft_reg_lists_bytes(d, lb, m, p);
    // partial spectra under 1 GiB, and within what a carved region has left
    const size_t fixed = Workspace::up(lists_bytes) + Workspace::up(chunk_bytes) + 256;
    const size_t budget = std::min<size_t>(size_t(1) << 30, ws.room() > fixed ? ws.room() - fixed : 0);
    nchunks = std::max<int64_t>(1, std::min<int64_t>(nchunks, int64_t(budget / chunk_bytes)));
    const int kper = n > 0 ? int((n + nchunks - 1) / nchunks) : 0;
    nchunks = n > 0 ? (n + kper - 1) / kper : 0;
    void *partial = nullptr, *V = nullptr, *lists = nullptr;
    if ((rc = ws.take(cur().partial, std::max<size_t>(1, size_t(nchunks) * chunk_bytes), &partial)) ||
        (rc = ws.take(cur().V, chunk_bytes, &V)) || (rc = ws.take(cur().entries, lists_bytes, &lists)))
      return rc;
    cm::RegJob j{};
    j.AT = static_cast<const float*>(AT);
    j.ld_at = ld_at;
    j.B = static_cast<const float*>(B);
    j.ld_b = ld_b;
    j.hA = sk->hA;
    j.hB = sk->hB;
    j.m = m;
    j.n = n;
    j.p = p;
    j.d = d;
    j.lb = lb;
    j.lbs = lbs;
    j.kper = kper;
    j.nchunks = int(nchunks);
    j.lists = lists;
    j.tw = twv;
    j.t2 = t2;
    j.scale = static_cast<const float2*>(psa.out);
    j.partial = static_cast<float*>(partial);
    const int f64_reduce = (flags & CM_FLAG_F64_REDUCE) || n > 16384;
    cm::ReduceArgs ra{fwht, 0, f64_reduce, partial, int(nchunks), d, lbv, cm::kRegLb2, twv, V};
    cm::CombineArgs ca{fwht, 0, V, d, lbv, cm::kRegLb2, twv, sk->P, lb};
    return run_stages(ws, {{"compress", [&] {
                              const hipError_t e = with_scales();
                              if (e != hipSuccess) return e;
                              if (grp) return fwht ? cm::launch_fwht_grp(j, cur().stream) : cm::launch_fft_grp(j, cur().stream);
                              return fwht ? cm::launch_fwht_reg(j, cur().stream) : cm::launch_fft_reg(j, cur().stream);
                            }},
                           {"reduce", [&] { return cm::launch_reduce_inverse(ra, cur().stream); }},
                           {"combine", [&] { return cm::launch_combine(ca, cur().stream); }}});
  }

  const int lb2 = std::min(lb, max_lb2);
  const int b1 = 1 << (lb - lb2);
  // whole rounds of the resident workgroups over all (t, slice, chunk)
  // units (a partial last round costs a whole unit's time), about 16 inner
  // indices per unit
  const int64_t per_chunk = int64_t(d) * b1;
  const int64_t slots = cm::compress_lds_slots(fwht, f64, lb2);
  const int64_t rounds = std::max<int64_t>(1, (int64_t(n) * per_chunk + slots * 16 - 1) / (slots * 16));
  int64_t nchunks = std::max<int64_t>(1, slots * rounds / per_chunk);
  nchunks = std::min<int64_t>(nchunks, std::max(1, n));
  const size_t chunk_bytes = size_t(d) * b * cplx * esz;
  // partial spectra under 1 GiB, and within what a carved region has left
  const size_t fixed = Workspace::up(chunk_bytes) + 256;
  const size_t cap = std::min<size_t>(size_t(1) << 30, ws.room() > fixed ? ws.room() - fixed : 0);
  nchunks = std::max<int64_t>(1, std::min<int64_t>(nchunks, int64_t(cap / chunk_bytes)));
  const int kper = n > 0 ? int((n + nchunks - 1) / nchunks) : 0;
  nchunks = n > 0 ? (n + kper - 1) / kper : 0;
  void *partial = nullptr, *V = nullptr;
  int rc;
  if ((rc = ws.take(cur().partial, std::max<size_t>(1, size_t(nchunks) * chunk_bytes), &partial)) ||
      (rc = ws.take(cur().V, chunk_bytes, &V)))
    return rc;
  cm::CompressArgs caa{};
  caa.fwht = fwht;
  caa.f64 = f64;
  caa.AT = AT;
  caa.ld_at = ld_at;
  caa.B = B;
  caa.ld_b = ld_b;
  caa.hA = sk->hA;
  caa.hB = sk->hB;
  caa.m = m;
  caa.n = n;
  caa.p = p;
  caa.d = d;
  caa.lb = lb;
  caa.lb2 = lb2;
  caa.nchunks = int(nchunks);
  caa.kper = kper;
  caa.tw = tw;
  caa.scale = psa.out;
  caa.partial = partial;
  const int f64_reduce = (flags & CM_FLAG_F64_REDUCE) || n > 16384;
  cm::ReduceArgs ra{fwht, f64, f64_reduce, partial, int(nchunks), d, lb, lb2, tw, V};
  cm::CombineArgs ca{fwht, f64, V, d, lb, lb2, tw, sk->P, lb};
  auto compress = [&] {
    const hipError_t e = with_scales();
    return e != hipSuccess ? e : cm::launch_compress_lds(caa, cur().stream);
  };
  auto reduce = [&] { return cm::launch_reduce_inverse(ra, cur().stream); };
  if (lb2 == lb) {  // one slice: the inverse writes the sketch, no combine
    ra.P = sk->P;
    return run_stages(ws, {{"compress", compress}, {"reduce", reduce}});
  }
  return run_stages(ws, {{"compress", compress},
                         {"reduce", reduce},
                         {"combine", [&] { return cm::launch_combine(ca, cur().stream); }}});
}

int compress_locked(int variant, int dtype, int m, int n, int p, const void* AT, int64_t ld_at,
                    const void* B, int64_t ld_b, int d, int b, uint64_t seed, int family,
                    int mem, unsigned flags, cm_sketch** inout) {
  int rc = init_locked(-1);
  if (rc) return rc;
  if (variant != CM_FFT && variant != CM_FWHT) return fail(CM_EINVAL, "variant must be CM_FFT or CM_FWHT");
  if (dtype != CM_F64 && dtype != CM_F32) return fail(CM_EINVAL, "dtype must be CM_F64 or CM_F32");
  if ((rc = check_db(d, b))) return rc;
  if ((rc = check_family(family))) return rc;
  if (m < 0 || n < 0 || p < 0) return fail(CM_EINVAL, "matrix dimensions must be non-negative");
  if (b > (1 << 30)) return fail(CM_ENOTSUP, "b above 2^30 is not supported");
  if (ld_at < m || ld_b < p) return fail(CM_EINVAL, "leading dimension smaller than the row length");
  if ((int64_t(n) * m > 0 && !AT) || (int64_t(n) * p > 0 && !B))
    return fail(CM_EINVAL, "null input pointer");
  if (mem != CM_HOST && mem != CM_DEVICE) return fail(CM_EINVAL, "mem must be CM_HOST or CM_DEVICE");
  const int fwht = variant == CM_FWHT, f64 = dtype == CM_F64;
  cm_sketch* sk = *inout;
  if (sk && (sk->fwht != fwht || sk->f64 != f64 || sk->d != d || sk->b != b || sk->m != m ||
             sk->p != p)) {
    return fail(CM_EINVAL, "sketch handle does not match (variant, dtype, d, b, m, p)");
  }
  bool fresh = false;
  if (!sk) {
    if ((rc = alloc_sketch(fwht, f64, d, b, m, p, &sk))) return rc;
    fresh = true;
  }
  sk->family = family;
  sk->seed = seed;
  const size_t esz = f64 ? 8 : 4;
  const void* dAT = AT;
  const void* dB = B;
  Workspace ws = call_workspace();  // the context's region, or the library's buffers
  if (mem == CM_HOST) {
    if ((rc = stage_in(ws, cur().AT, AT, n, m, ld_at, esz, &dAT)) ||
        (rc = stage_in(ws, cur().B, B, n, p, ld_b, esz, &dB))) {
      if (fresh) cm_sketch_free(sk);
      return rc;
    }
  }
  if ((rc = build_hashes(family, d, sk->lb, seed, m, p, sk->hA, sk->hB, ws)) ||
      (rc = compress_device(sk, dAT, ld_at, dB, ld_b, n, flags, ws))) {
    if (fresh) cm_sketch_free(sk);
    return rc;
  }
  *inout = sk;
  return CM_OK;
}

// Device buffers of a recovery pass (in the context's workspace or the
// library's buffers).
struct RecoverBufs {
  unsigned long long* counter = nullptr;
  int32_t* oi = nullptr;
  int32_t* oj = nullptr;
  void* ov = nullptr;
};

int decompress_locked(const cm_sketch* sk, void* C, int64_t ldc, int mem, int recover, double tau,
                      int64_t max_out, RecoverBufs* rec = nullptr) {
  Workspace ws = call_workspace();
  cm::DecompressArgs da{};
  da.fwht = sk->fwht;
  da.f64 = sk->f64;
  da.P = sk->P;
  da.hA = sk->hA;
  da.hB = sk->hB;
  da.m = sk->m;
  da.p = sk->p;
  da.d = sk->d;
  da.b = sk->b;
  da.ldc = ldc;
  const size_t esz = sk->f64 ? 8 : 4;
  int rc;
  void* dC = nullptr;
  if (C && mem == CM_HOST) {
    // the device copy is dense (pitch p); the copy back writes only the m x p
    // window of the caller's matrix, so columns p .. ldc-1 stay untouched
    if ((rc = ws.take(cur().C, std::max<size_t>(1, size_t(sk->m) * sk->p * esz), &dC))) return rc;
    da.C = dC;
    da.ldc = sk->p;
  } else {
    da.C = C;
  }
  if (recover) {
    void *cnt = nullptr, *oi = nullptr, *oj = nullptr, *ov = nullptr;
    if ((rc = ws.take(cur().counter, 8, &cnt)) ||
        (rc = ws.take(cur().oi, std::max<int64_t>(1, max_out) * 4, &oi)) ||
        (rc = ws.take(cur().oj, std::max<int64_t>(1, max_out) * 4, &oj)) ||
        (rc = ws.take(cur().ov, std::max<int64_t>(1, max_out) * esz, &ov)))
      return rc;
    CM_HIP(hipMemsetAsync(cnt, 0, 8, cur().stream));
    da.recover = 1;
    da.tau = tau;
    da.counter = static_cast<unsigned long long*>(cnt);
    da.max_out = max_out;
    da.oi = static_cast<int32_t*>(oi);
    da.oj = static_cast<int32_t*>(oj);
    da.ov = ov;
    if (rec) *rec = RecoverBufs{da.counter, da.oi, da.oj, da.ov};
  }
  if ((rc = timed("decompress", [&] { return cm::launch_decompress(da, cur().stream); }))) return rc;
  if (C && mem == CM_HOST && sk->m > 0 && sk->p > 0) {
    // a 2-D copy into pageable host memory moves one row at a time (about
    // 3 ms for 512 rows), so the rows travel as one span: into a contiguous
    // C through the pinned transfer slots, else through the pinned staging
    // buffer and placed on the host (a carved workspace copies directly)
    const size_t row = size_t(sk->p) * esz, span = row * size_t(sk->m);
    if (ldc == sk->p && !ws.carved()) {
      if ((rc = d2h(C, dC, span))) return rc;
    } else if (ldc == sk->p) {
      CM_HIP(hipMemcpyAsync(C, dC, span, hipMemcpyDeviceToHost, cur().stream));
      CM_HIP(hipStreamSynchronize(cur().stream));
    } else if (!ws.carved()) {
      if ((rc = ensure_pinned(span))) return rc;
      CM_HIP(hipMemcpyAsync(cur().pinned, dC, span, hipMemcpyDeviceToHost, cur().stream));
      CM_HIP(hipStreamSynchronize(cur().stream));
      const char* src = static_cast<const char*>(cur().pinned);
      char* dst = static_cast<char*>(C);
      for (int i = 0; i < sk->m; ++i) std::memcpy(dst + size_t(i) * size_t(ldc) * esz, src + size_t(i) * row, row);
    } else {
      CM_HIP(hipMemcpy2DAsync(C, size_t(ldc) * esz, dC, row, row, size_t(sk->m), hipMemcpyDeviceToHost,
                              cur().stream));
      CM_HIP(hipStreamSynchronize(cur().stream));
    }
  }
  return CM_OK;
}

// cm_matmul_noalloc: compress and full decompression of device operands with
// every intermediate (hash arrays, balance scales, twiddle tables, scatter
// schedules, partial spectra, inverse, sketch) carved from the fixed device
// scratch of noalloc.hip: no device or host allocation, no upload.
int matmul_noalloc_locked(int variant, int dtype, int m, int n, int p, const void* AT,
                          int64_t ld_at, const void* B, int64_t ld_b, void* C, int64_t ldc,
                          int d, int b, uint64_t seed, int family) {
  int rc = init_locked(-1);
  if (rc) return rc;
  if (variant != CM_FFT && variant != CM_FWHT) return fail(CM_EINVAL, "variant must be CM_FFT or CM_FWHT");
  if (dtype != CM_F64 && dtype != CM_F32) return fail(CM_EINVAL, "dtype must be CM_F64 or CM_F32");
  if ((rc = check_db(d, b)) || (rc = check_family(family))) return rc;
  if (m < 0 || n < 0 || p < 0) return fail(CM_EINVAL, "matrix dimensions must be non-negative");
  if (b > (1 << 30)) return fail(CM_ENOTSUP, "b above 2^30 is not supported");
  if (ld_at < m || ld_b < p || ldc < p)
    return fail(CM_EINVAL, "leading dimension smaller than the row length");
  if ((int64_t(n) * m > 0 && !AT) || (int64_t(n) * p > 0 && !B) || (int64_t(m) * p > 0 && !C))
    return fail(CM_EINVAL, "null operand pointer");
  void* base = nullptr;
  size_t cap = 0;
  if (cm::noalloc_scratch(cur().device, &base, &cap))
    return fail(CM_EDEVICE, "no-allocation scratch is not set up on this device");
  Workspace ws;
  ws.base = static_cast<char*>(base);
  ws.cap = cap;
  const int fwht = variant == CM_FWHT, f64 = dtype == CM_F64;
  const size_t esz = f64 ? 8 : 4;
  cm_sketch sk{};
  sk.fwht = fwht;
  sk.f64 = f64;
  sk.d = d;
  sk.b = b;
  sk.lb = ilog2(b);
  sk.m = m;
  sk.p = p;
  sk.family = family;
  sk.seed = seed;
  DevBuf none;
  void *hA = nullptr, *hB = nullptr, *P = nullptr;
  if ((rc = ws.take(none, std::max<size_t>(4, size_t(d) * m * 4), &hA)) ||
      (rc = ws.take(none, std::max<size_t>(4, size_t(d) * p * 4), &hB)) ||
      (rc = ws.take(none, size_t(d) * b * esz, &P)))
    return rc;
  sk.hA = static_cast<uint32_t*>(hA);
  sk.hB = static_cast<uint32_t*>(hB);
  sk.P = P;
  if ((rc = build_hashes(family, d, sk.lb, seed, m, p, sk.hA, sk.hB, ws, true)) ||
      (rc = compress_device(&sk, AT, ld_at, B, ld_b, n, 0, ws)))
    return rc;
  cm::DecompressArgs da{};
  da.fwht = fwht;
  da.f64 = f64;
  da.P = sk.P;
  da.hA = sk.hA;
  da.hB = sk.hB;
  da.m = m;
  da.p = p;
  da.d = d;
  da.b = b;
  da.C = C;
  da.ldc = ldc;
  CM_HIP(cm::launch_decompress(da, cur().stream));
  CM_HIP(hipStreamSynchronize(cur().stream));
  return CM_OK;
}

int matmul_dropin(int variant, int n_a, int n_inner, int n_b, double* C, const double* AT,
                  const double* B, int d, int b, uint64_t seed, int seed_given, int family) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int rc = check_db(d, b);
  if (rc) return rc;
  if ((rc = check_family(family))) return rc;
  if (!seed_given) seed = cm::time_seed();
  cm_sketch* sk = nullptr;
  rc = compress_locked(variant, CM_F64, n_a, n_inner, n_b, AT, n_a, B, n_b, d, b, seed, family,
                       CM_HOST, 0, &sk);
  if (rc) return rc;
  rc = decompress_locked(sk, C, n_b, CM_HOST, 0, 0.0, 0);
  cm_sketch_free(sk);
  return rc;
}

// The recovered entries of a sketch, sorted by (i, j), on the current context.
int recover_locked(const cm_sketch* sk, double tau, int64_t max_out, int32_t* out_i, int32_t* out_j,
                   double* out_val, int64_t* count) {
  RecoverBufs rb;
  int rc = decompress_locked(sk, nullptr, sk->p, CM_DEVICE, 1, tau, max_out, &rb);
  if (rc) return rc;
  unsigned long long found = 0;
  CM_HIP(hipMemcpyAsync(&found, rb.counter, 8, hipMemcpyDeviceToHost, cur().stream));
  CM_HIP(hipStreamSynchronize(cur().stream));
  *count = int64_t(found);
  const int64_t k = std::min<int64_t>(int64_t(found), max_out);
  if (k == 0) return CM_OK;
  std::vector<int32_t> vi(k), vj(k);
  std::vector<double> vv(k);
  CM_HIP(hipMemcpyAsync(vi.data(), rb.oi, k * 4, hipMemcpyDeviceToHost, cur().stream));
  CM_HIP(hipMemcpyAsync(vj.data(), rb.oj, k * 4, hipMemcpyDeviceToHost, cur().stream));
  if (sk->f64) {
    CM_HIP(hipMemcpyAsync(vv.data(), rb.ov, k * 8, hipMemcpyDeviceToHost, cur().stream));
    CM_HIP(hipStreamSynchronize(cur().stream));
  } else {
    std::vector<float> vf(k);
    CM_HIP(hipMemcpyAsync(vf.data(), rb.ov, k * 4, hipMemcpyDeviceToHost, cur().stream));
    CM_HIP(hipStreamSynchronize(cur().stream));
    for (int64_t q = 0; q < k; ++q) vv[q] = vf[q];
  }
  cm::sort_recovered(k, sk->p, vi.data(), vj.data(), vv.data());
  for (int64_t q = 0; q < k; ++q) {
    if (out_i) out_i[q] = vi[q];
    if (out_j) out_j[q] = vj[q];
    if (out_val) out_val[q] = vv[q];
  }
  return CM_OK;
}

// Wait until the kernels that wrote a sketch are done: its context's stream
// is the default one's, or a context's (then the whole of its device).
int sketch_ready(const cm_sketch* sk) {
  if (sk->on_default) {
    CM_HIP(hipStreamSynchronize(g_default.stream));
    return CM_OK;
  }
  int dev = -1;
  CM_HIP(hipGetDevice(&dev));
  if (dev != sk->device) CM_HIP(hipSetDevice(sk->device));
  const hipError_t e = hipDeviceSynchronize();
  if (dev != sk->device) CM_HIP(hipSetDevice(dev));
  CM_HIP(e);
  return CM_OK;
}

// compress with the handle's reuse rules, synchronous for host operands
int compress_entry(int variant, int dtype, int n_a, int n_inner, int n_b, const void* AT,
                   int64_t ld_at, const void* B, int64_t ld_b, int d, int b, uint64_t seed,
                   int family, int mem, unsigned flags, cm_sketch** inout) {
  int rc = compress_locked(variant, dtype, n_a, n_inner, n_b, AT, ld_at, B, ld_b, d, b, seed,
                           family, mem, flags, inout);
  if (!rc && mem == CM_HOST) CM_HIP(hipStreamSynchronize(cur().stream));
  if (!rc) (*inout)->on_default = g_cur == &g_default;
  return rc;
}

int check_decompress_args(const cm_sketch* sk, void* C, int64_t ldc, int mem) {
  if (!sk) return fail(CM_EINVAL, "null sketch");
  if (!C && int64_t(sk->m) * sk->p > 0) return fail(CM_EINVAL, "null output pointer");
  if (ldc < sk->p) return fail(CM_EINVAL, "ldc smaller than the number of columns");
  if (mem != CM_HOST && mem != CM_DEVICE) return fail(CM_EINVAL, "mem must be CM_HOST or CM_DEVICE");
  return CM_OK;
}

// every buffer, table, event and (if owned) the stream of a context's state
void release_state(State& st, bool own_stream) {
  if (st.stream) (void)hipStreamSynchronize(st.stream);
  for (DevBuf* b : {&st.partial, &st.V, &st.params, &st.tables, &st.AT, &st.B, &st.C, &st.counter,
                    &st.oi, &st.oj, &st.ov, &st.plainA, &st.plainB, &st.entries, &st.scale}) {
    if (b->p) (void)hipFree(b->p);
    *b = DevBuf{};
  }
  if (st.fft_t2) (void)hipFree(st.fft_t2);
  for (auto& kv : st.twiddles) (void)hipFree(kv.second);
  st.twiddles.clear();
  if (st.pinned) (void)hipHostFree(st.pinned);
  if (st.xfer) (void)hipHostFree(st.xfer);
  for (hipEvent_t e : st.xfer_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& r : st.pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : st.event_pool) (void)hipEventDestroy(e);
  if (st.ws_base) (void)hipFree(st.ws_base);
  if (own_stream && st.stream) (void)hipStreamDestroy(st.stream);
  st = State{};
}

}  // namespace

// A context: its own state (device, stream, buffers or workspace), and a
// lock that serializes the calls made on it.
struct cm_context {
  State st;
  std::recursive_mutex mu;
  bool own_stream = false;
};

namespace {

// The scope of one call on a context: its lock, its state as the thread's
// current context, and its device (the thread's device is restored after).
struct CtxCall {
  std::lock_guard<std::recursive_mutex> lk;
  State* prev;
  int prev_dev = -1;
  explicit CtxCall(cm_context* c) : lk(c->mu), prev(g_cur) {
    if (hipGetDevice(&prev_dev) != hipSuccess) prev_dev = -1;
    g_cur = &c->st;
    if (prev_dev != c->st.device) (void)hipSetDevice(c->st.device);
  }
  ~CtxCall() {
    g_cur = prev;
    int now = -1;
    if (prev_dev >= 0 && hipGetDevice(&now) == hipSuccess && now != prev_dev) (void)hipSetDevice(prev_dev);
  }
  CtxCall(const CtxCall&) = delete;
  CtxCall& operator=(const CtxCall&) = delete;
};

}  // namespace

namespace cm {
namespace rt {
std::recursive_mutex& mutex() { return g_mu; }
int init() { return init_locked(-1); }
hipStream_t stream() { return cur().stream; }
int error(int code, const std::string& msg) { return fail(code, msg); }
int twiddles(int lb, int f64, const void** out) { return ::twiddles(lb, f64, out); }
int fft_t2_table(const void** out) { return ::fft_t2_table(out); }
int timed(const char* name, const std::function<hipError_t()>& launch) {
  return ::timed(name, launch);
}
int timed_on(hipStream_t stream, const char* name, const std::function<hipError_t()>& launch) {
  return ::timed_on(stream, name, launch);
}
}  // namespace rt
}  // namespace cm

extern "C" {

const char* cm_last_error(void) { return g_err.c_str(); }

int cm_init(int device) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return init_locked(device);
}

uint64_t cm_default_seed(void) { return cm::time_seed(); }

int cm_matmul_fft(int n_a, int n_inner, int n_b, double* C, const double* AT, const double* B,
                  int d, int b, uint64_t seed, int seed_given, int family) {
  return matmul_dropin(CM_FFT, n_a, n_inner, n_b, C, AT, B, d, b, seed, seed_given, family);
}

int cm_matmul_fwht(int n_a, int n_inner, int n_b, double* C, const double* AT, const double* B,
                   int d, int b, uint64_t seed, int seed_given, int family) {
  return matmul_dropin(CM_FWHT, n_a, n_inner, n_b, C, AT, B, d, b, seed, seed_given, family);
}

int cm_matmul_noalloc(int variant, int dtype, int n_a, int n_inner, int n_b, const void* AT,
                      int64_t ld_at, const void* B, int64_t ld_b, void* C, int64_t ldc, int d,
                      int b, uint64_t seed, int family) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return matmul_noalloc_locked(variant, dtype, n_a, n_inner, n_b, AT, ld_at, B, ld_b, C, ldc, d, b,
                               seed, family);
}

int cm_gemm(int n_a, int n_inner, int n_b, double* C, const double* AT, const double* B) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int rc = init_locked(-1);
  if (rc) return rc;
  if (n_a < 0 || n_inner < 0 || n_b < 0) return fail(CM_EINVAL, "matrix dimensions must be non-negative");
  const void *dAT, *dB;
  Workspace lib;  // the library's buffers
  if ((rc = stage_in(lib, cur().AT, AT, n_inner, n_a, n_a, 8, &dAT)) ||
      (rc = stage_in(lib, cur().B, B, n_inner, n_b, n_b, 8, &dB)))
    return rc;
  if ((rc = ensure_buf(cur().C, std::max<size_t>(8, size_t(n_a) * n_b * 8)))) return rc;
  if (n_inner == 0) {
    CM_HIP(hipMemsetAsync(cur().C.p, 0, size_t(n_a) * n_b * 8, cur().stream));
  } else if ((rc = timed("gemm", [&] {
                return cm::launch_gemm_f64(n_a, n_inner, n_b, static_cast<const double*>(dAT),
                                           static_cast<const double*>(dB),
                                           static_cast<double*>(cur().C.p), cur().stream);
              }))) {
    return rc;
  }
  if (size_t(n_a) * n_b && (rc = d2h(C, cur().C.p, size_t(n_a) * n_b * 8))) return rc;
  CM_HIP(hipStreamSynchronize(cur().stream));
  return CM_OK;
}

int cm_hash_arrays(int family, int d, int b, uint64_t seed, int m, int p, uint32_t* outA,
                   uint32_t* outB, int mem) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int rc = init_locked(-1);
  if (rc) return rc;
  if ((rc = check_db(d, b)) || (rc = check_family(family))) return rc;
  if (m < 0 || p < 0) return fail(CM_EINVAL, "key counts must be non-negative");
  uint32_t *dA = outA, *dB = outB;
  DevBuf tA, tB;
  if (mem == CM_HOST) {
    CM_HIP(hipMalloc(&tA.p, std::max<size_t>(4, size_t(d) * m * 4)));
    CM_HIP(hipMalloc(&tB.p, std::max<size_t>(4, size_t(d) * p * 4)));
    dA = static_cast<uint32_t*>(tA.p);
    dB = static_cast<uint32_t*>(tB.p);
  }
  Workspace lib;  // the library's buffers
  rc = build_hashes(family, d, ilog2(b), seed, m, p, dA, dB, lib);
  if (!rc && mem == CM_HOST) {
    if (size_t(d) * m) CM_HIP(hipMemcpyAsync(outA, dA, size_t(d) * m * 4, hipMemcpyDeviceToHost, cur().stream));
    if (size_t(d) * p) CM_HIP(hipMemcpyAsync(outB, dB, size_t(d) * p * 4, hipMemcpyDeviceToHost, cur().stream));
  }
  (void)hipStreamSynchronize(cur().stream);
  if (tA.p) (void)hipFree(tA.p);
  if (tB.p) (void)hipFree(tB.p);
  return rc;
}

int cm_compress(int variant, int dtype, int n_a, int n_inner, int n_b, const void* AT,
                int64_t ld_at, const void* B, int64_t ld_b, int d, int b, uint64_t seed,
                int family, int mem, cm_sketch** out) {
  if (!out) return fail(CM_EINVAL, "null output handle");
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  *out = nullptr;
  return compress_entry(variant, dtype, n_a, n_inner, n_b, AT, ld_at, B, ld_b, d, b, seed, family,
                        mem, 0, out);
}

int cm_compress_ex(int variant, int dtype, int n_a, int n_inner, int n_b, const void* AT,
                   int64_t ld_at, const void* B, int64_t ld_b, int d, int b, uint64_t seed,
                   int family, int mem, unsigned flags, cm_sketch** inout) {
  if (!inout) return fail(CM_EINVAL, "null handle pointer");
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return compress_entry(variant, dtype, n_a, n_inner, n_b, AT, ld_at, B, ld_b, d, b, seed, family,
                        mem, flags, inout);
}

int cm_decompress(const cm_sketch* sk, void* C, int64_t ldc, int mem) {
  int rc = check_decompress_args(sk, C, ldc, mem);
  if (rc) return rc;
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return decompress_locked(sk, C, ldc, mem, 0, 0.0, 0);
}

int cm_recover(const cm_sketch* sk, double tau, int64_t max_out, int32_t* out_i, int32_t* out_j,
               double* out_val, int64_t* count) {
  if (!sk || !count) return fail(CM_EINVAL, "null argument");
  if (max_out < 0) return fail(CM_EINVAL, "max_out must be non-negative");
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return recover_locked(sk, tau, max_out, out_i, out_j, out_val, count);
}

int cm_sketch_values(const cm_sketch* sk, double* host_out) {
  if (!sk || !host_out) return fail(CM_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  const size_t n = size_t(sk->d) * sk->b;
  if (int rc = sketch_ready(sk)) return rc;
  if (sk->f64) {
    CM_HIP(hipMemcpy(host_out, sk->P, n * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(n);
    CM_HIP(hipMemcpy(tmp.data(), sk->P, n * 4, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < n; ++q) host_out[q] = tmp[q];
  }
  return CM_OK;
}

int cm_sketch_hashes(const cm_sketch* sk, uint32_t* outA, uint32_t* outB) {
  if (!sk) return fail(CM_EINVAL, "null sketch");
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = sketch_ready(sk)) return rc;
  if (outA && size_t(sk->d) * sk->m)
    CM_HIP(hipMemcpy(outA, sk->hA, size_t(sk->d) * sk->m * 4, hipMemcpyDeviceToHost));
  if (outB && size_t(sk->d) * sk->p)
    CM_HIP(hipMemcpy(outB, sk->hB, size_t(sk->d) * sk->p * 4, hipMemcpyDeviceToHost));
  return CM_OK;
}

void cm_sketch_free(cm_sketch* sk) {
  if (!sk) return;
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  (void)sketch_ready(sk);
  if (sk->hA) (void)hipFree(sk->hA);
  if (sk->hB) (void)hipFree(sk->hB);
  if (sk->P) (void)hipFree(sk->P);
  delete sk;
}

int cm_dev_alloc(size_t bytes, void** ptr) {
  if (!ptr) return fail(CM_EINVAL, "null pointer");
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int rc = init_locked(-1);
  if (rc) return rc;
  CM_HIP(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
  return CM_OK;
}

int cm_dev_free(void* ptr) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (cur().stream) (void)hipStreamSynchronize(cur().stream);
  if (ptr) CM_HIP(hipFree(ptr));
  return CM_OK;
}

int cm_host_alloc(size_t bytes, void** ptr) {
  if (!ptr) return fail(CM_EINVAL, "null pointer");
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int rc = init_locked(-1);
  if (rc) return rc;
  CM_HIP(hipHostMalloc(ptr, std::max<size_t>(bytes, 1), hipHostMallocDefault));
  return CM_OK;
}

int cm_host_free(void* ptr) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (ptr) CM_HIP(hipHostFree(ptr));
  return CM_OK;
}

int cm_memcpy(void* dst, const void* src, size_t bytes, int dst_mem, int src_mem) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int rc = init_locked(-1);
  if (rc) return rc;
  if (bytes == 0) return CM_OK;
  hipMemcpyKind kind = dst_mem == CM_DEVICE ? (src_mem == CM_DEVICE ? hipMemcpyDeviceToDevice
                                                                    : hipMemcpyHostToDevice)
                                            : (src_mem == CM_DEVICE ? hipMemcpyDeviceToHost
                                                                    : hipMemcpyHostToHost);
  CM_HIP(hipMemcpyAsync(dst, src, bytes, kind, cur().stream));
  CM_HIP(hipStreamSynchronize(cur().stream));
  return CM_OK;
}

int cm_synchronize(void) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int rc = init_locked(-1);
  if (rc) return rc;
  CM_HIP(hipStreamSynchronize(cur().stream));
  return CM_OK;
}

int cm_timing_enable(int on) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  resolve_timing();
  cur().timing = on != 0;
  return CM_OK;
}

int cm_timing_only(const char* name) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  resolve_timing();
  cur().timing_only = name ? name : "";
  return CM_OK;
}

int cm_timing_reset(void) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  resolve_timing();
  cur().totals.clear();
  return CM_OK;
}

int cm_timing_get(const char* name, double* total_ms, int64_t* launches) {
  if (!name) return fail(CM_EINVAL, "null name");
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  resolve_timing();
  auto it = cur().totals.find(name);
  if (total_ms) *total_ms = it == cur().totals.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == cur().totals.end() ? 0 : it->second.second;
  return CM_OK;
}

// ---------------------------------------------------------------- contexts

int cm_context_create(int device, void* stream, size_t workspace_bytes, cm_context** out) {
  if (!out) return fail(CM_EINVAL, "null output handle");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(CM_EDEVICE, "no HIP device is available (compmatmul_amd needs an MI355X)");
  if (device < 0 || device >= count) return fail(CM_EINVAL, "device index out of range");
  hipDeviceProp_t prop;
  CM_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(CM_EDEVICE, std::string("compmatmul_amd is built for gfx950, found ") + prop.gcnArchName);
  auto ctx = std::make_unique<cm_context>();
  ctx->st.device = device;
  int rc = CM_OK;
  {
    CtxCall call(ctx.get());  // the new context's device, restored after
    if (stream) {
      ctx->st.stream = static_cast<hipStream_t>(stream);
    } else if (hipStreamCreateWithFlags(&ctx->st.stream, hipStreamNonBlocking) != hipSuccess) {
      ctx->st.stream = nullptr;
      rc = fail(CM_EDEVICE, "stream creation failed");
    } else {
      ctx->own_stream = true;
    }
    if (!rc && workspace_bytes) {
      void* p = nullptr;
      const hipError_t e = hipMalloc(&p, workspace_bytes);
      if (e != hipSuccess) {
        rc = fail(e == hipErrorOutOfMemory ? CM_ENOMEM : CM_EDEVICE,
                  "context workspace of " + std::to_string(workspace_bytes) + " bytes: " + hipGetErrorString(e));
      } else {
        ctx->st.ws_base = static_cast<char*>(p);
        ctx->st.ws_cap = workspace_bytes;
      }
    }
    if (rc) release_state(ctx->st, ctx->own_stream);
  }
  if (rc) return rc;
  *out = ctx.release();
  return CM_OK;
}

int cm_context_destroy(cm_context* ctx) {
  if (!ctx) return CM_OK;
  {
    CtxCall call(ctx);
    release_state(ctx->st, ctx->own_stream);
  }
  delete ctx;
  return CM_OK;
}

int cm_ctx_compress(cm_context* ctx, int variant, int dtype, int n_a, int n_inner, int n_b,
                    const void* AT, int64_t ld_at, const void* B, int64_t ld_b, int d, int b,
                    uint64_t seed, int family, int mem, unsigned flags, cm_sketch** inout) {
  if (!ctx || !inout) return fail(CM_EINVAL, "null argument");
  CtxCall call(ctx);
  return compress_entry(variant, dtype, n_a, n_inner, n_b, AT, ld_at, B, ld_b, d, b, seed, family,
                        mem, flags, inout);
}

int cm_ctx_decompress(cm_context* ctx, const cm_sketch* sk, void* C, int64_t ldc, int mem) {
  if (!ctx) return fail(CM_EINVAL, "null context");
  int rc = check_decompress_args(sk, C, ldc, mem);
  if (rc) return rc;
  CtxCall call(ctx);
  return decompress_locked(sk, C, ldc, mem, 0, 0.0, 0);
}

int cm_ctx_recover(cm_context* ctx, const cm_sketch* sk, double tau, int64_t max_out, int32_t* out_i,
                   int32_t* out_j, double* out_val, int64_t* count) {
  if (!ctx || !sk || !count) return fail(CM_EINVAL, "null argument");
  if (max_out < 0) return fail(CM_EINVAL, "max_out must be non-negative");
  CtxCall call(ctx);
  return recover_locked(sk, tau, max_out, out_i, out_j, out_val, count);
}

int cm_ctx_synchronize(cm_context* ctx) {
  if (!ctx) return fail(CM_EINVAL, "null context");
  CtxCall call(ctx);
  CM_HIP(hipStreamSynchronize(ctx->st.stream));
  return CM_OK;
}

}  // extern "C"
