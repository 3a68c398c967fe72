// Parallel host copies (see host_copy.hpp).
#include "host_copy.hpp"

#include <pthread.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>

namespace cm {

namespace {

constexpr size_t kMinParallel = size_t(512) << 10;  // smaller copies: the caller alone
constexpr int kMaxWorkers = 3;                      // plus the caller: four parts

class CopyPool {
 public:
  explicit CopyPool(int workers) : workers_(workers) {
    for (int i = 1; i <= workers_; ++i) std::thread([this, i] { run(i); }).detach();
  }

  // false: the workers are busy with another caller's copy
  bool copy(void* dst, const void* src, size_t len) {
    std::unique_lock<std::mutex> owner(busy_, std::try_to_lock);
    if (!owner.owns_lock()) return false;
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = static_cast<char*>(dst);
      src_ = static_cast<const char*>(src);
      len_ = len;
      pending_ = workers_;
      ++gen_;
    }
    work_cv_.notify_all();
    part(0, dst_, src_, len);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return pending_ == 0; });
    return true;
  }

 private:
  // part i of workers_ + 1, split at 64-byte boundaries
  void part(int i, char* dst, const char* src, size_t len) const {
    const size_t parts = size_t(workers_) + 1;
    const size_t per = (len / parts + 63) & ~size_t(63);
    const size_t lo = std::min(len, per * size_t(i));
    const size_t hi = i + 1 == int(parts) ? len : std::min(len, lo + per);
    if (hi > lo) std::memcpy(dst + lo, src + lo, hi - lo);
  }

  void run(int id) {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(mu_);
      work_cv_.wait(lk, [&] { return gen_ != seen; });
      seen = gen_;
      char* dst = dst_;
      const char* src = src_;
      const size_t len = len_;
      lk.unlock();
      part(id, dst, src, len);
      lk.lock();
      if (--pending_ == 0) done_cv_.notify_one();
    }
  }

  const int workers_;
  std::mutex busy_;  // held by the caller whose copy the workers run
  std::mutex mu_;
  std::condition_variable work_cv_, done_cv_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t len_ = 0;
  uint64_t gen_ = 0;
  int pending_ = 0;
};

// The pool is created on first use and never destroyed: its threads wait
// on it until the process ends.  A forked child starts without threads, so
// it drops the parent's pool (and creates its own on first use); the pool
// lock is held across fork so the child never inherits it locked by
// another thread.
std::mutex g_pool_mu;
CopyPool* g_pool = nullptr;

void fork_prepare() { g_pool_mu.lock(); }
void fork_parent() { g_pool_mu.unlock(); }
void fork_child() {
  g_pool = nullptr;
  g_pool_mu.unlock();
}

CopyPool* pool() {
  static std::once_flag atfork;
  std::call_once(atfork, [] { pthread_atfork(fork_prepare, fork_parent, fork_child); });
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (!g_pool) {
    const unsigned hw = std::thread::hardware_concurrency();
    const int workers = std::min<int>(kMaxWorkers, hw > 1 ? int(hw) - 1 : 0);
    g_pool = new CopyPool(workers);
  }
  return g_pool;
}

}  // namespace

void parallel_copy(void* dst, const void* src, size_t len) {
  if (len < kMinParallel || !pool()->copy(dst, src, len)) std::memcpy(dst, src, len);
}

}  // namespace cm
