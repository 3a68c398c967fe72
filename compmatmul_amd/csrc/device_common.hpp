// Device-side helpers shared by the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cm {

// Claims the next item of a work queue whose counter is `c` and which holds
// `len` items: returns the item's index, or a value >= len once the queue
// is exhausted.  With `probe`, a relaxed load first passes over an exhausted
// queue without an atomic: failed claims by every resident workgroup on one
// counter serialize in L2 with the successful ones.  The load costs a round
// trip of its own, so callers probe only where the queue may be exhausted
// (see probe_next).
__device__ __forceinline__ unsigned claim_item(unsigned* c, int len, bool probe) {
  if (probe) {
    const unsigned seen = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (int(seen) >= len) return seen;
  }
  return atomicAdd(c, 1u);
}

// Whether the next claim on a queue of `len` items should probe: on a queue
// taken over from another XCD (qi > 0) or one no longer than the grid's
// share of it, and once the last item claimed (`last`, -1 before any) leaves
// no more than that share.  Measured on MI355X: probing every claim slowed
// the 512-product batch by 4%; probing only a workgroup's first claim on
// each queue left one failed atomic per workgroup and up to twice the
// compress time on small products.
__device__ __forceinline__ bool probe_next(int qi, int len, int last) {
  const int share = int(gridDim.x >> 3) + 1;
  return qi > 0 || len - last <= share;
}

template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> csqr(cpx<T> a) {
  return {a.x * a.x - a.y * a.y, T(2) * a.x * a.y};
}

// Packed hash entry: bucket in bits 0..30, bit 31 set when the sign is -1.
__device__ __forceinline__ uint32_t hbucket(uint32_t e) { return e & 0x7fffffffu; }
__device__ __forceinline__ uint32_t hneg(uint32_t e) { return e >> 31; }

template <typename T>
__device__ __forceinline__ T apply_sign(T v, uint32_t neg) { return neg ? -v : v; }

// Schedule placement of one 64-key chunk (one key per lane) whose keys
// fall into rounds 0..7 (r < 0: the key is not in this list): returns the
// slot of this lane's key, run[r] plus the keys of round r in lower lanes,
// and advances every run[rr] by its round's count in the chunk.  Three
// ballots on the bits of r and one on validity give every lane the mask of
// its round; the runs stay wave-uniform.
__device__ __forceinline__ int64_t round_slot(int r, int lane, int64_t (&run)[8]) {
  const bool ok = r >= 0;
  const uint64_t bv = __ballot(ok);
  const uint64_t b0 = __ballot(ok && (r & 1));
  const uint64_t b1 = __ballot(ok && (r & 2));
  const uint64_t b2 = __ballot(ok && (r & 4));
  uint64_t same = bv & ((r & 1) ? b0 : ~b0) & ((r & 2) ? b1 : ~b1) & ((r & 4) ? b2 : ~b2);
  const int below = __popcll(same & ((1ull << lane) - 1ull));
  int64_t mine = run[0];
#pragma unroll
  for (int rr = 1; rr < 8; ++rr) mine = r == rr ? run[rr] : mine;
#pragma unroll
  for (int rr = 0; rr < 8; ++rr)
    run[rr] += __popcll(bv & ((rr & 1) ? b0 : ~b0) & ((rr & 2) ? b1 : ~b1) & ((rr & 4) ? b2 : ~b2));
  return mine + below;
}

// reverse the low `bits` bits of x (bits >= 1)
__device__ __forceinline__ uint32_t bitrev(uint32_t x, int bits) {
  return __brev(x) >> (32 - bits);
}

}  // namespace cm
