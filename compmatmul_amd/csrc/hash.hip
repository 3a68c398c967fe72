// Device evaluation of the drawn hash functions into packed bucket/sign arrays.
//
// Per-key functions follow hashing.hpp bit for bit:
//   multiply-shift      bucket (a*x) >> (64 - log2 b), sign = bit 63  (:91-97)
//   multiply-add-shift  bucket (a*x + b) >> (64 - log2 b), sign = bit 63 (:163-169)
//   tabulation          h = XOR_c T[c][char c of x]; bucket h >> (32 - log2 b),
//                       sign = bit 31 (:298-314)
// A set top bit means sign +1; the packed entry stores bit 31 = 1 for -1.
#include "device_common.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

__device__ __forceinline__ uint32_t tab_raw(const uint32_t* T, int family, int tab_len,
                                            uint32_t x) {
  if (family == 2) {
    return T[0 * tab_len + (x & 0xffu)] ^ T[1 * tab_len + ((x >> 8) & 0xffu)] ^
           T[2 * tab_len + ((x >> 16) & 0xffu)] ^ T[3 * tab_len + (x >> 24)];
  }
  return T[0 * tab_len + (x & 0xffffu)] ^ T[1 * tab_len + (x >> 16)];
}

__device__ __forceinline__ uint64_t prm(const HashArgs& a, int idx) {
  return a.use_inline ? a.inline_params[idx] : a.mul_params[idx];
}

// bucket and "is negative" of function pair (h, s) at key x
__device__ __forceinline__ uint32_t eval_packed(const HashArgs& a, int fh, int fs, uint32_t x) {
  uint32_t bucket, neg;
  if (a.family == 0 || a.family == 1) {
    const int w = a.family == 0 ? 2 : 1;
    const uint64_t ah = prm(a, fh * w), as = prm(a, fs * w);
    const uint64_t bh = w == 2 ? prm(a, fh * w + 1) : 0ull;
    const uint64_t bs = w == 2 ? prm(a, fs * w + 1) : 0ull;
    const uint64_t vh = ah * uint64_t(x) + bh;
    const uint64_t vs = as * uint64_t(x) + bs;
    bucket = uint32_t(vh >> (64 - a.lb));
    neg = uint32_t(vs >> 63) ^ 1u;
  } else {
    const int chars = a.family == 2 ? 4 : 2;
    const uint32_t* th = a.tables + int64_t(fh) * chars * a.tab_len;
    const uint32_t* ts = a.tables + int64_t(fs) * chars * a.tab_len;
    bucket = tab_raw(th, a.family, a.tab_len, x) >> (32 - a.lb);
    neg = (tab_raw(ts, a.family, a.tab_len, x) >> 31) ^ 1u;
  }
  return bucket | (neg << 31);
}

// functions are stored h1[0..d), h2[0..d), s1[0..d), s2[0..d)
__device__ __forceinline__ void hash_body(const HashArgs& a, int blk) {
  const int64_t idx = int64_t(blk) * 256 + threadIdx.x;
  const int64_t na = int64_t(a.d) * a.m;
  const int64_t total = na + int64_t(a.d) * a.p;
  if (idx >= total) return;
  if (idx < na) {
    const int t = int(idx / a.m);
    const uint32_t key = uint32_t(idx % a.m);
    a.outA[idx] = eval_packed(a, t, 2 * a.d + t, key);
  } else {
    const int64_t r = idx - na;
    const int t = int(r / a.p);
    const uint32_t key = uint32_t(r % a.p);
    a.outB[r] = eval_packed(a, a.d + t, 3 * a.d + t, key);
  }
}

// The parameter block is read in place from the kernel-argument segment (it
// is the only argument, at offset 0): indexing the by-value argument with a
// run-time index would copy all of it (2 KB) into every lane's scratch.
__global__ void __launch_bounds__(256) k_hash(HashArgs) {
  const void* seg = (const void*)__builtin_amdgcn_kernarg_segment_ptr();
  hash_body(*static_cast<const HashArgs*>(seg), blockIdx.x);
}

__global__ void __launch_bounds__(256) k_hash_batch(const HashArgs* __restrict__ jobs,
                                                    const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  hash_body(jobs[e.x], e.y);
}

}  // namespace

int hash_batch_blocks(const HashArgs& a) {
  return int((int64_t(a.d) * (int64_t(a.m) + a.p) + 255) / 256);
}

hipError_t launch_hash_batch(const HashArgs* jobs, const int2* blocks, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  k_hash_batch<<<dim3(unsigned(nblocks)), dim3(256), 0, s>>>(jobs, blocks);
  return hipGetLastError();
}

hipError_t launch_hash(const HashArgs& a, hipStream_t s) {
  const int64_t total = int64_t(a.d) * (int64_t(a.m) + a.p);
  if (total == 0) return hipSuccess;
  const int64_t blocks = (total + 255) / 256;
  k_hash<<<dim3(unsigned(blocks)), dim3(256), 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace cm
