// Request server: continuous batching of compressed products on one GPU
// (cm_server_*).
//
// One dispatcher thread owns the device work.  Each pass it
//   1. retires finished waves: reads each request's recovered list (and C)
//      back, stamps its completion time, and returns the wave's memory;
//   2. if a wave slot is free, admits queued requests in arrival order while
//      their footprint fits the HBM pool (at most max_wave_products), uploads
//      their host operands on the slot's stream (device-resident operands are
//      read in place; CSR operands are expanded on the device), plans them as
//      one batch in the pool and launches it.
// Up to max_waves waves run at once on separate streams, so the persistent
// compress kernels of a new wave fill the CUs that an older wave's tail
// leaves idle.  Memory is one device allocation made at creation and carved
// first-fit; a request holds two blocks: its wave's share (operands,
// schedules, partial spectra, recovery list) until the wave retires, and its
// hash arrays and sketch until the client releases it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/compmatmul_amd.h"
#include "batch_internal.hpp"
#include "recover_sort.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace {

using cm::rt::error;

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int ilog2(int64_t b) {
  int l = 0;
  while ((int64_t(1) << l) < b) ++l;
  return l;
}

// first-fit allocator over [0, size) with coalescing of free neighbours
class Pool {
 public:
  explicit Pool(size_t size) { free_[0] = size; }
  bool alloc(size_t bytes, size_t* off) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t start = it->first + bytes;
      free_.erase(it);
      if (rest) free_[start] = rest;
      used_ += bytes;
      sizes_[*off] = bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off) {
    auto s = sizes_.find(off);
    if (s == sizes_.end()) return;
    size_t start = off, len = s->second;
    used_ -= len;
    sizes_.erase(s);
    auto next = free_.lower_bound(start);
    if (next != free_.end() && start + len == next->first) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        prev->second += len;
        return;
      }
    }
    free_[start] = len;
  }
  size_t largest() const {
    size_t m = 0;
    for (auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }
  size_t peak() const { return peak_; }

 private:
  std::map<size_t, size_t> free_;   // offset -> length
  std::map<size_t, size_t> sizes_;  // allocated offset -> length
  size_t used_ = 0, peak_ = 0;
};

enum State { kQueued, kRunning, kDone, kFailed };

struct Request {
  uint64_t ticket = 0;
  cm_request r{};
  State state = kQueued;
  int rc = CM_OK;
  std::string err;
  bool release_when_done = false;
  double t_submit = 0, t_done = 0;
  size_t keep_bytes = 0, wave_bytes = 0;  // footprint estimates
  size_t keep_off = 0;
  bool has_keep = false;
  cm::Retained keep{};
  std::vector<int32_t> oi, oj;
  std::vector<float> ov;
  int64_t found = 0;
};

// Pinned host memory of a slot: the planner's uploads (reused once the
// slot's wave has retired, so its stream has passed them), the wave's
// recovery counters (copied on the wave's stream before its completion
// event) and the recovered lists read back at retirement.
constexpr size_t kStagingBytes = size_t(16) << 20;

struct Slot {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Request*> reqs;
  std::vector<float*> dev_C;  // per request, device C or nullptr
  cm_batch* bt = nullptr;
  size_t off = 0;
  char* staging = nullptr;                // kStagingBytes
  unsigned long long* counts = nullptr;   // max_wave recovery counters
  char* lists = nullptr;                  // recovered lists, grown on demand
  size_t lists_bytes = 0;
};

}  // namespace

struct cm_server {
  size_t budget = 0;
  int max_waves = 2, max_wave = 64, grain = 64;
  char* base = nullptr;
  std::unique_ptr<Pool> pool;
  std::mutex mu;  // requests, queue, pool, stats
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_ticket = 1;
  std::vector<Slot> slots;
  std::thread thread;
  int device = 0;  // the library's device at creation; the dispatcher thread selects it
  bool stop = false;
  cm_server_stats stats{};
};

namespace {

cm_product as_product(const cm_request& r) {
  cm_product q{};
  q.variant = r.variant;
  q.m = r.m;
  q.n = r.n;
  q.p = r.p;
  q.ld_at = r.m;
  q.ld_b = r.p;
  q.ldc = r.p;
  q.d = r.d;
  q.b = r.b;
  q.seed = r.seed;
  q.family = r.family;
  q.tau = r.tau;
  q.max_out = std::isnan(r.tau) ? 0 : r.max_out;
  return q;
}

size_t keep_bytes(const cm_request& r) {
  return align_up(size_t(r.d) * r.m * 4) + align_up(size_t(r.d) * r.p * 4) +
         align_up(size_t(r.d) * r.b * 4);
}

size_t nnz(const int32_t* indptr, int n) { return indptr ? size_t(indptr[n]) : 0; }

// operand staging + dense operands + C of one request, in its wave block
// (device-resident dense operands and C are used in place)
size_t operand_bytes(const cm_request& r) {
  const bool dev = r.mem == CM_DEVICE;
  size_t s = 0;
  if (!dev || r.format == CM_CSR) s += align_up(size_t(r.n) * r.m * 4) + align_up(size_t(r.n) * r.p * 4);
  if (r.format == CM_CSR && !dev) {
    s += 2 * align_up(size_t(r.n + 1) * 4);
    s += 2 * align_up(nnz(r.at_indptr, r.n) * 4) + 2 * align_up(nnz(r.b_indptr, r.n) * 4);
  }
  if (r.C && !dev) s += align_up(size_t(r.m) * r.p * 4);
  return s;
}

int check_csr(const int32_t* indptr, const int32_t* indices, const float* values, int rows,
              int cols, const char* name) {
  if (!indptr || (indptr[rows] > 0 && (!indices || !values)))
    return error(CM_EINVAL, std::string(name) + ": missing CSR arrays");
  if (indptr[0] != 0) return error(CM_EINVAL, std::string(name) + ": indptr[0] must be 0");
  for (int r = 0; r < rows; ++r)
    if (indptr[r + 1] < indptr[r])
      return error(CM_EINVAL, std::string(name) + ": indptr must be non-decreasing");
  for (int64_t k = 0; k < indptr[rows]; ++k)
    if (indices[k] < 0 || indices[k] >= cols)
      return error(CM_EINVAL, std::string(name) + ": column index out of range");
  return CM_OK;
}

void finish_request(cm_server* sv, Request* q, int rc, const std::string& err) {
  // caller holds sv->mu
  q->rc = rc;
  q->err = err;
  q->state = rc == CM_OK ? kDone : kFailed;
  q->t_done = now_s();
  if (rc == CM_OK)
    ++sv->stats.requests_done;
  else
    ++sv->stats.requests_failed;
  if (q->has_keep && (rc != CM_OK || q->release_when_done)) {
    sv->pool->release(q->keep_off);
    q->has_keep = false;
  }
  if (q->release_when_done) sv->reqs.erase(q->ticket);
}

// read back and retire a finished wave (dispatcher thread): the counters
// are already in the slot's pinned memory; every list of the wave is copied
// behind them with one wait
void retire(cm_server* sv, Slot& w) {
  const double t0 = now_s();
  std::vector<int> rcs(w.reqs.size(), CM_OK);
  std::vector<std::string> errs(w.reqs.size());
  {
    std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
    std::vector<int64_t> keep(w.reqs.size(), 0), at(w.reqs.size(), 0);
    size_t need = 0;
    for (size_t k = 0; k < w.reqs.size(); ++k) {
      Request* q = w.reqs[k];
      if (std::isnan(q->r.tau)) continue;
      q->found = int64_t(w.counts[k]);
      keep[k] = std::min<int64_t>(q->found, q->r.max_out);
      at[k] = int64_t(need);
      need += size_t(keep[k]) * 12;
    }
    hipError_t e = hipSuccess;
    if (need > w.lists_bytes) {
      if (w.lists) (void)hipHostFree(w.lists);
      w.lists = nullptr;
      w.lists_bytes = 0;
      e = hipHostMalloc(reinterpret_cast<void**>(&w.lists), std::max<size_t>(need, size_t(1) << 20),
                        hipHostMallocDefault);
      if (e == hipSuccess) w.lists_bytes = std::max<size_t>(need, size_t(1) << 20);
    }
    bool copies = false;
    for (size_t k = 0; k < w.reqs.size() && e == hipSuccess; ++k) {
      if (keep[k] == 0) continue;
      const cm::ProductView v = cm::batch_product(w.bt, int(k));
      char* dst = w.lists + at[k];
      e = hipMemcpyAsync(dst, v.oi, keep[k] * 4, hipMemcpyDeviceToHost, w.s);
      if (e == hipSuccess) e = hipMemcpyAsync(dst + keep[k] * 4, v.oj, keep[k] * 4, hipMemcpyDeviceToHost, w.s);
      if (e == hipSuccess) e = hipMemcpyAsync(dst + keep[k] * 8, v.ov, keep[k] * 4, hipMemcpyDeviceToHost, w.s);
      copies = true;
    }
    if (e == hipSuccess && copies) e = hipStreamSynchronize(w.s);
    for (size_t k = 0; k < w.reqs.size(); ++k) {
      Request* q = w.reqs[k];
      if (!std::isnan(q->r.tau)) {
        if (e != hipSuccess) {
          rcs[k] = CM_EDEVICE;
          errs[k] = std::string("reading results: ") + hipGetErrorString(e);
          continue;
        }
        const int64_t kk = keep[k];
        q->oi.resize(kk);
        q->oj.resize(kk);
        q->ov.resize(kk);
        if (kk > 0) {
          const char* src = w.lists + at[k];
          std::memcpy(q->oi.data(), src, kk * 4);
          std::memcpy(q->oj.data(), src + kk * 4, kk * 4);
          std::memcpy(q->ov.data(), src + kk * 8, kk * 4);
        }
        cm::sort_recovered(kk, q->r.p, q->oi.data(), q->oj.data(), q->ov.data());
      }
      if (q->r.C && w.dev_C[k] && q->r.m > 0 && q->r.p > 0) {
        hipError_t e = hipMemcpyAsync(q->r.C, w.dev_C[k], size_t(q->r.m) * q->r.p * 4,
                                      hipMemcpyDeviceToHost, w.s);
        if (e == hipSuccess) e = hipStreamSynchronize(w.s);
        if (e != hipSuccess) {
          rcs[k] = CM_EDEVICE;
          errs[k] = std::string("reading C: ") + hipGetErrorString(e);
        }
      }
    }
    cm::batch_destroy(w.bt);
    w.bt = nullptr;
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->stats.retire_s += now_s() - t0;
  for (size_t k = 0; k < w.reqs.size(); ++k) finish_request(sv, w.reqs[k], rcs[k], errs[k]);
  sv->pool->release(w.off);
  w.reqs.clear();
  w.dev_C.clear();
  w.busy = false;
  sv->cv_done.notify_all();
}

// admit, upload, plan and launch one wave on slot w (dispatcher thread);
// returns false when nothing was launched
bool launch_wave(cm_server* sv, Slot& w) {
  const double t0 = now_s();
  std::vector<Request*> take;
  size_t wave_off = 0, wave_bytes = 0;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->queue.empty()) return false;
    size_t want = 0;
    for (Request* q : sv->queue) {
      if (int(take.size()) >= sv->max_wave) break;
      const size_t need = q->keep_bytes + q->wave_bytes;
      if (want + need > sv->pool->largest()) break;
      want += need;
      take.push_back(q);
    }
    if (take.empty()) {
      Request* head = sv->queue.front();
      if (head->keep_bytes + head->wave_bytes > sv->budget) {
        sv->queue.pop_front();
        finish_request(sv, head, CM_ENOMEM, "request footprint exceeds the server's HBM budget");
        sv->cv_done.notify_all();
      }
      return false;  // wait for memory to free up
    }
    // retained blocks, then one block for the wave
    size_t k = 0;
    for (; k < take.size(); ++k) {
      Request* q = take[k];
      if (!sv->pool->alloc(q->keep_bytes, &q->keep_off)) break;
      q->has_keep = true;
      char* p = sv->base + q->keep_off;
      q->keep.hA = reinterpret_cast<uint32_t*>(p);
      p += align_up(size_t(q->r.d) * q->r.m * 4);
      q->keep.hB = reinterpret_cast<uint32_t*>(p);
      p += align_up(size_t(q->r.d) * q->r.p * 4);
      q->keep.P = reinterpret_cast<float*>(p);
    }
    for (size_t u = k; u < take.size(); ++u) take[u]->has_keep = false;
    take.resize(k);
    while (!take.empty()) {
      wave_bytes = 0;
      for (Request* q : take) wave_bytes += q->wave_bytes;
      if (sv->pool->alloc(wave_bytes, &wave_off)) break;
      Request* q = take.back();  // shrink the wave until its block fits
      sv->pool->release(q->keep_off);
      q->has_keep = false;
      take.pop_back();
    }
    if (take.empty()) return false;
    for (size_t u = 0; u < take.size(); ++u) {
      sv->queue.pop_front();
      take[u]->state = kRunning;
    }
  }

  // uploads and the batch, on the slot's stream
  std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
  char* cur = sv->base + wave_off;
  auto carve = [&](size_t bytes) {
    char* p = cur;
    cur += align_up(bytes);
    return p;
  };
  std::vector<cm_product> prods(take.size());
  std::vector<cm::Retained> keep(take.size());
  w.dev_C.assign(take.size(), nullptr);
  int rc = CM_OK;
  std::string err;
  for (size_t k = 0; k < take.size() && rc == CM_OK; ++k) {
    const cm_request& r = take[k]->r;
    const bool dev = r.mem == CM_DEVICE;
    cm_product q = as_product(r);
    hipError_t e = hipSuccess;
    float* dAT = nullptr;
    float* dB = nullptr;
    if (dev && r.format == CM_DENSE) {  // resident operands, read in place
      dAT = const_cast<float*>(r.AT);
      dB = const_cast<float*>(r.B);
    } else {
      dAT = reinterpret_cast<float*>(carve(size_t(r.n) * r.m * 4));
      dB = reinterpret_cast<float*>(carve(size_t(r.n) * r.p * 4));
    }
    if (dev) {
      if (r.format == CM_CSR) {  // resident CSR arrays, expanded in place
        e = cm::launch_csr_dense(r.at_indptr, r.at_indices, r.at_values, r.n, r.m, r.m, dAT, w.s);
        if (e == hipSuccess)
          e = cm::launch_csr_dense(r.b_indptr, r.b_indices, r.b_values, r.n, r.p, r.p, dB, w.s);
      }
    } else if (r.format == CM_DENSE) {
      if (size_t(r.n) * r.m) e = hipMemcpyAsync(dAT, r.AT, size_t(r.n) * r.m * 4, hipMemcpyHostToDevice, w.s);
      if (e == hipSuccess && size_t(r.n) * r.p)
        e = hipMemcpyAsync(dB, r.B, size_t(r.n) * r.p * 4, hipMemcpyHostToDevice, w.s);
    } else {
      struct Csr {
        const int32_t *ip, *ix;
        const float* v;
        int cols;
        float* out;
      } ops[2] = {{r.at_indptr, r.at_indices, r.at_values, r.m, dAT},
                  {r.b_indptr, r.b_indices, r.b_values, r.p, dB}};
      for (const Csr& c : ops) {
        const size_t z = nnz(c.ip, r.n);
        int32_t* ip = reinterpret_cast<int32_t*>(carve(size_t(r.n + 1) * 4));
        int32_t* ix = reinterpret_cast<int32_t*>(carve(z * 4));
        float* v = reinterpret_cast<float*>(carve(z * 4));
        if (e == hipSuccess) e = hipMemcpyAsync(ip, c.ip, size_t(r.n + 1) * 4, hipMemcpyHostToDevice, w.s);
        if (e == hipSuccess && z) e = hipMemcpyAsync(ix, c.ix, z * 4, hipMemcpyHostToDevice, w.s);
        if (e == hipSuccess && z) e = hipMemcpyAsync(v, c.v, z * 4, hipMemcpyHostToDevice, w.s);
        if (e == hipSuccess) e = cm::launch_csr_dense(ip, ix, v, r.n, c.cols, c.cols, c.out, w.s);
      }
    }
    if (e != hipSuccess) {
      rc = CM_EDEVICE;
      err = std::string("operand upload: ") + hipGetErrorString(e);
      break;
    }
    q.AT = dAT;
    q.B = dB;
    if (r.C && dev) {
      q.C = r.C;  // written in place, nothing to read back
    } else if (r.C) {
      w.dev_C[k] = reinterpret_cast<float*>(carve(size_t(r.m) * r.p * 4));
      q.C = w.dev_C[k];
    }
    prods[k] = q;
    keep[k] = take[k]->keep;
  }
  if (rc == CM_OK) {
    const size_t used = size_t(cur - (sv->base + wave_off));
    cm::PlanEnv env;
    env.stream = w.s;
    env.arena = cur;
    env.arena_bytes = wave_bytes > used ? wave_bytes - used : 0;
    env.keep = keep.data();
    env.staging = w.staging;
    env.staging_bytes = kStagingBytes;
    rc = cm::batch_plan_env(int(prods.size()), prods.data(), 0, sv->grain, env, &w.bt);
    if (rc == CM_OK) rc = cm::batch_enqueue(w.bt);
    // the recovery counters follow the wave on its stream, so they are on
    // the host when its completion event fires
    if (rc == CM_OK && hipMemcpyAsync(w.counts, cm::batch_counters(w.bt), prods.size() * 8,
                                      hipMemcpyDeviceToHost, w.s) != hipSuccess)
      rc = error(CM_EDEVICE, "counter read-back failed");
    if (rc == CM_OK && hipEventRecord(w.done, w.s) != hipSuccess) rc = error(CM_EDEVICE, "event record failed");
    if (rc) err = cm_last_error();
  }
  if (rc != CM_OK) {
    (void)hipStreamSynchronize(w.s);
    if (w.bt) cm::batch_destroy(w.bt);
    w.bt = nullptr;
    std::lock_guard<std::mutex> lk2(sv->mu);
    for (Request* q : take) finish_request(sv, q, rc, err);
    sv->pool->release(wave_off);
    sv->cv_done.notify_all();
    return true;
  }
  w.reqs = take;
  w.off = wave_off;
  w.busy = true;
  std::lock_guard<std::mutex> lk2(sv->mu);
  sv->stats.launch_s += now_s() - t0;
  ++sv->stats.waves;
  int inflight = 0;
  for (const Slot& s : sv->slots) inflight += s.busy;
  sv->stats.max_waves_in_flight = std::max(sv->stats.max_waves_in_flight, inflight);
  return true;
}

void dispatcher(cm_server* sv) {
  (void)hipSetDevice(sv->device);  // the device the pool and streams were made on
  for (;;) {
    bool progress = false;
    for (Slot& w : sv->slots) {
      if (w.busy && hipEventQuery(w.done) == hipSuccess) {
        retire(sv, w);
        progress = true;
      }
    }
    bool stopping;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      stopping = sv->stop;
    }
    if (!stopping) {
      for (Slot& w : sv->slots)
        if (!w.busy && launch_wave(sv, w)) progress = true;
    }
    bool any_busy = false;
    for (const Slot& w : sv->slots) any_busy |= w.busy;
    if (stopping && !any_busy) break;
    if (!progress) {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait_for(lk, std::chrono::microseconds(any_busy ? 50 : 2000));
    }
  }
  std::lock_guard<std::mutex> lk(sv->mu);
  while (!sv->queue.empty()) {
    finish_request(sv, sv->queue.front(), CM_EINVAL, "server stopped");
    sv->queue.pop_front();
  }
  sv->cv_done.notify_all();
}

// a slot's stream (after its work), event and pinned buffers
void release_slot(Slot& w) {
  if (w.s) {
    (void)hipStreamSynchronize(w.s);
    (void)hipStreamDestroy(w.s);
  }
  if (w.done) (void)hipEventDestroy(w.done);
  if (w.staging) (void)hipHostFree(w.staging);
  if (w.counts) (void)hipHostFree(w.counts);
  if (w.lists) (void)hipHostFree(w.lists);
  w = Slot{};
}

Request* find(cm_server* sv, uint64_t ticket) {
  auto it = sv->reqs.find(ticket);
  return it == sv->reqs.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

int cm_server_create(size_t hbm_bytes, int max_waves, int max_wave_products, int grain,
                     cm_server** out) {
  if (!out) return error(CM_EINVAL, "null output handle");
  // HIP gives a process four hardware queues by default: more wave streams
  // would share them and not run concurrently
  if (hbm_bytes < (size_t(1) << 20) || max_waves < 1 || max_waves > 4 || max_wave_products < 1 ||
      grain < 1)
    return error(CM_EINVAL, "need hbm_bytes >= 1 MiB, 1 <= max_waves <= 4, max_wave_products >= 1, grain >= 1");
  std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
  int rc = cm::rt::init();
  if (rc) return rc;
  auto sv = std::make_unique<cm_server>();
  if (hipGetDevice(&sv->device) != hipSuccess) return error(CM_EDEVICE, "no current device");
  sv->budget = align_up(hbm_bytes);
  sv->max_waves = max_waves;
  sv->max_wave = max_wave_products;
  sv->grain = grain;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&sv->base), sv->budget);
  if (e != hipSuccess)
    return error(e == hipErrorOutOfMemory ? CM_ENOMEM : CM_EDEVICE,
                 std::string("server pool: ") + hipGetErrorString(e));
  sv->pool = std::make_unique<Pool>(sv->budget);
  sv->slots.resize(max_waves);
  for (Slot& w : sv->slots) {
    if (hipStreamCreateWithFlags(&w.s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&w.done, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&w.staging), kStagingBytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&w.counts), size_t(max_wave_products) * 8,
                      hipHostMallocDefault) != hipSuccess) {
      for (Slot& x : sv->slots) release_slot(x);
      (void)hipFree(sv->base);
      return error(CM_EDEVICE, "stream, event or pinned staging creation failed");
    }
  }
  sv->stats.hbm_bytes = sv->budget;
  cm_server* raw = sv.release();
  raw->thread = std::thread(dispatcher, raw);
  *out = raw;
  return CM_OK;
}

int cm_server_submit(cm_server* sv, const cm_request* r, uint64_t* ticket) {
  if (!sv || !r || !ticket) return error(CM_EINVAL, "null argument");
  if (r->variant != CM_FFT && r->variant != CM_FWHT) return error(CM_EINVAL, "variant must be CM_FFT or CM_FWHT");
  if (r->m < 0 || r->n < 0 || r->p < 0) return error(CM_EINVAL, "negative dimension");
  if (r->b < 2 || (r->b & (r->b - 1)) != 0) return error(CM_EINVAL, "b must be a power of two at least 2");
  if (r->d < 1 || r->d > 16) return error(CM_ENOTSUP, "requests take 1 <= d <= 16");
  if (r->family < CM_MULTIPLY_ADD_SHIFT || r->family > CM_TABULATION16) return error(CM_EHASH, "unknown hash family");
  const int lb = ilog2(r->b);
  if (!(r->variant == CM_FWHT ? cm::fwht_reg_supported(lb, r->m, r->p) : cm::fft_reg_supported(lb, r->m, r->p)))
    return error(CM_ENOTSUP, "shape outside the register kernels' range");
  if (!std::isnan(r->tau) && r->max_out < 0) return error(CM_EINVAL, "max_out must be >= 0");
  int rc;
  if (r->mem != CM_HOST && r->mem != CM_DEVICE) return error(CM_EINVAL, "mem must be CM_HOST or CM_DEVICE");
  if (r->format == CM_DENSE) {
    if (r->n > 0 && ((r->m > 0 && !r->AT) || (r->p > 0 && !r->B))) return error(CM_EINVAL, "missing operand");
  } else if (r->format == CM_CSR && r->mem == CM_DEVICE) {
    if (r->n > 0 && (!r->at_indptr || !r->b_indptr)) return error(CM_EINVAL, "missing CSR arrays");
  } else if (r->format == CM_CSR) {
    if ((rc = check_csr(r->at_indptr, r->at_indices, r->at_values, r->n, r->m, "AT")) ||
        (rc = check_csr(r->b_indptr, r->b_indices, r->b_values, r->n, r->p, "B")))
      return rc;
  } else {
    return error(CM_EINVAL, "format must be CM_DENSE or CM_CSR");
  }
  auto q = std::make_unique<Request>();
  q->r = *r;
  q->keep_bytes = keep_bytes(*r);
  size_t arena = 0;
  {
    std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
    cm_product p = as_product(*r);
    p.AT = reinterpret_cast<const float*>(kAlign);  // placeholders: sizing only
    p.B = reinterpret_cast<const float*>(kAlign);
    cm::Retained none{};
    cm::PlanEnv env;
    env.keep = &none;
    env.size_only = &arena;
    int grain;
    {
      std::lock_guard<std::mutex> lk2(sv->mu);
      grain = sv->grain;
    }
    if ((rc = cm::batch_plan_env(1, &p, 0, grain, env, nullptr))) return rc;
  }
  // a wave's tables and alignment add a little per product
  q->wave_bytes = operand_bytes(*r) + arena + 4 * kAlign + 8192;
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sv->stop) return error(CM_EINVAL, "server stopped");
  q->ticket = sv->next_ticket++;
  q->t_submit = now_s();
  *ticket = q->ticket;
  sv->queue.push_back(q.get());
  sv->reqs[q->ticket] = std::move(q);
  sv->cv_work.notify_one();
  return CM_OK;
}

int cm_server_wait(cm_server* sv, uint64_t ticket, double timeout_s) {
  if (!sv) return error(CM_EINVAL, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto ready = [&] {
    Request* q = find(sv, ticket);
    return !q || q->state == kDone || q->state == kFailed;
  };
  if (timeout_s < 0)
    sv->cv_done.wait(lk, ready);
  else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return 1;
  Request* q = find(sv, ticket);
  if (!q) return error(CM_EINVAL, "unknown ticket");
  if (q->state == kFailed) return error(q->rc, q->err);
  return CM_OK;
}

int cm_server_result(cm_server* sv, uint64_t ticket, int64_t max_out, int32_t* out_i,
                     int32_t* out_j, float* out_val, int64_t* count, double* latency_s) {
  if (!sv) return error(CM_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  Request* q = find(sv, ticket);
  if (!q) return error(CM_EINVAL, "unknown ticket");
  if (q->state == kFailed) return error(q->rc, q->err);
  if (q->state != kDone) return error(CM_EINVAL, "request not done");
  if (count) *count = q->found;
  if (latency_s) *latency_s = q->t_done - q->t_submit;
  const int64_t k = std::min<int64_t>(std::max<int64_t>(max_out, 0), int64_t(q->oi.size()));
  for (int64_t t = 0; t < k; ++t) {
    if (out_i) out_i[t] = q->oi[t];
    if (out_j) out_j[t] = q->oj[t];
    if (out_val) out_val[t] = q->ov[t];
  }
  return CM_OK;
}

int cm_server_decompress(cm_server* sv, uint64_t ticket, float* host_C) {
  if (!sv || !host_C) return error(CM_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
  cm::DecompressArgs da{};
  size_t c_off = 0;
  {
    std::lock_guard<std::mutex> lk2(sv->mu);
    Request* q = find(sv, ticket);
    if (!q) return error(CM_EINVAL, "unknown ticket");
    if (q->state != kDone || !q->has_keep) return error(CM_EINVAL, "request not done or its sketch released");
    da.fwht = q->r.variant == CM_FWHT;
    da.P = q->keep.P;
    da.hA = q->keep.hA;
    da.hB = q->keep.hB;
    da.m = q->r.m;
    da.p = q->r.p;
    da.d = q->r.d;
    da.b = q->r.b;
    da.ldc = q->r.p;
    if (size_t(da.m) * da.p == 0) return CM_OK;
    if (!sv->pool->alloc(size_t(da.m) * da.p * 4, &c_off)) return error(CM_ENOMEM, "no room for C in the server's HBM budget");
  }
  da.C = sv->base + c_off;
  hipStream_t s = cm::rt::stream();
  hipError_t e = cm::launch_decompress(da, s);
  if (e == hipSuccess) e = hipMemcpyAsync(host_C, da.C, size_t(da.m) * da.p * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  std::lock_guard<std::mutex> lk2(sv->mu);
  sv->pool->release(c_off);
  return e == hipSuccess ? CM_OK : error(CM_EDEVICE, std::string("decompress: ") + hipGetErrorString(e));
}

int cm_server_values(cm_server* sv, uint64_t ticket, float* host_out) {
  if (!sv || !host_out) return error(CM_EINVAL, "null argument");
  std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
  const float* P = nullptr;
  size_t bytes = 0;
  {
    std::lock_guard<std::mutex> lk2(sv->mu);
    Request* q = find(sv, ticket);
    if (!q) return error(CM_EINVAL, "unknown ticket");
    if (q->state != kDone || !q->has_keep) return error(CM_EINVAL, "request not done or its sketch released");
    P = q->keep.P;
    bytes = size_t(q->r.d) * size_t(q->r.b) * sizeof(float);
  }
  hipStream_t s = cm::rt::stream();
  hipError_t e = hipMemcpyAsync(host_out, P, bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? CM_OK : error(CM_EDEVICE, std::string("reading the sketch: ") + hipGetErrorString(e));
}

int cm_server_release(cm_server* sv, uint64_t ticket) {
  if (!sv) return error(CM_EINVAL, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  Request* q = find(sv, ticket);
  if (!q) return error(CM_EINVAL, "unknown ticket");
  if (q->state == kQueued || q->state == kRunning) {
    q->release_when_done = true;
    return CM_OK;
  }
  if (q->has_keep) sv->pool->release(q->keep_off);
  sv->reqs.erase(ticket);
  sv->cv_work.notify_one();
  return CM_OK;
}

int cm_server_stats_get(cm_server* sv, cm_server_stats* out) {
  if (!sv || !out) return error(CM_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  *out = sv->stats;
  out->peak_bytes = sv->pool->peak();
  return CM_OK;
}

int cm_server_destroy(cm_server* sv) {
  if (!sv) return CM_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  if (sv->thread.joinable()) sv->thread.join();
  {
    std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
    for (Slot& w : sv->slots) release_slot(w);
    if (sv->base) (void)hipFree(sv->base);
  }
  delete sv;
  return CM_OK;
}

}  // extern "C"
