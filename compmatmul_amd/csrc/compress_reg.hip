// Register-resident single-wave FP32 FWHT compress (the fast path).
//
// Same mathematics as the LDS kernel in compress.hip (sliced, modulated
// count-sketch into b2 = 4096 buckets, natural-order Walsh-Hadamard
// transform, pointwise product, accumulation over a chunk of k), laid out
// for one 64-lane wavefront per workgroup:
//
//  * A slice of 4096 values lives in 64 VGPRs per lane: element
//    e = r*64 + lane after the sketch is read from LDS.  The WHT is a tensor
//    product of 2-point butterflies over the 12 index bits, so the six
//    register bits are transformed in registers, the 64x64 block is
//    transposed through LDS, and the six lane
//    bits -- now register bits -- are transformed in registers.  After that
//    lane L, register r holds frequency f = L*64 + r.  The transpose uses a
//    row stride of 65 floats: conflict-free, and every LDS address is one
//    per-lane base plus an immediate offset.
//  * The spectral accumulator (64 VGPRs) stays in registers for the whole
//    chunk of k; only the final sum is written.
//  * The scatter uses per-(t, slice) precomputed entries (LDS byte offset |
//    combined sign bit), so each nonzero costs two ANDs, one XOR and one LDS
//    float add.  One wave owns its LDS buffer, so the scatter order -- and
//    the result -- is deterministic.
//  * One 16.25 KB LDS buffer per wave serves the scatter, the read-back and
//    both transposes; eight waves fit a CU by LDS.
//  * Workgroups are numbered so that the (t, slice) units of one chunk of k
//    run on the same XCD and share its L2 copy of the rows of AT and B.
#include "device_common.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

constexpr int LB2 = 12;           // slice length 4096
constexpr int B2 = 1 << LB2;
constexpr int R = B2 / 64;        // values per lane
constexpr int SROW = 65;          // padded row stride of the transpose
constexpr int SLEN = 64 * SROW;   // LDS floats per wave (4160)

__device__ __forceinline__ float flip(float v, uint32_t sgn) {
  return __int_as_float(__float_as_int(v) ^ sgn);
}

// packed scatter entries: entry = (lo * 4) | (sign << 31), where lo is the
// bucket modulo b2 and sign combines s(i) with the slice modulation
// (-1)^popc((h >> log2 b2) & slice)
__global__ void __launch_bounds__(256) k_reg_entries(const uint32_t* __restrict__ hA,
                                                     const uint32_t* __restrict__ hB, int m,
                                                     int p, int d, int lbs,
                                                     uint32_t* __restrict__ EA,
                                                     uint32_t* __restrict__ EB) {
  const int b1 = 1 << lbs;
  const int64_t na = int64_t(d) * b1 * m;
  const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (idx >= na + int64_t(d) * b1 * p) return;
  const bool sb = idx >= na;
  const int64_t r = sb ? idx - na : idx;
  const int keys = sb ? p : m;
  const int key = int(r % keys);
  const int64_t ts = r / keys;  // t * b1 + slice
  const int slice = int(ts & (b1 - 1));
  const int t = int(ts >> lbs);
  const uint32_t e = (sb ? hB : hA)[int64_t(t) * keys + key];
  const uint32_t h = hbucket(e);
  const uint32_t neg = hneg(e) ^ (__popc((h >> LB2) & uint32_t(slice)) & 1u);
  (sb ? EB : EA)[r] = ((h & (B2 - 1)) << 2) | (neg << 31);
}

__device__ __forceinline__ void wht_reg(float (&x)[R]) {
#pragma unroll
  for (int s = 1; s < R; s <<= 1) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (r & s) continue;
      const float a = x[r], b = x[r | s];
      x[r] = a + b;
      x[r | s] = a - b;
    }
  }
}

__device__ __forceinline__ void zero_lds(float* S, int lane) {
  float4* S4 = reinterpret_cast<float4*>(S);
#pragma unroll
  for (int j = 0; j < SLEN / 4 / 64; ++j) S4[j * 64 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
  if (lane < (SLEN / 4) % 64) S4[(SLEN / 4 / 64) * 64 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// sketch one row into S (zeroed), read it into x, zero S, full WHT of x.
// The row and its entries are loaded in batches of 32 per lane so that the
// loads are in flight together before the LDS adds consume them.
constexpr int BATCH = 32;

__device__ __forceinline__ void sketch_transform(const float* __restrict__ row,
                                                 const uint32_t* __restrict__ ent, int keys,
                                                 float* S, int lane, float (&x)[R]) {
  char* Sb = reinterpret_cast<char*>(S);
  for (int base = 0; base < keys; base += 64 * BATCH) {
    float v[BATCH];
    uint32_t e[BATCH];
#pragma unroll
    for (int q = 0; q < BATCH; ++q) {
      // lanes past the end re-read the last key and add -0.0 to S[0], an
      // exact no-op, so the loop needs no branches
      const int i = base + q * 64 + lane;
      const int ic = min(i, keys - 1);
      const float vq = row[ic];
      const uint32_t eq = ent[ic];
      v[q] = i < keys ? vq : -0.f;
      e[q] = i < keys ? eq : 0u;
    }
#pragma unroll
    for (int q = 0; q < BATCH; ++q)
      __hip_atomic_fetch_add(reinterpret_cast<float*>(Sb + (e[q] & 0xffffu)),
                             flip(v[q], e[q] & 0x80000000u), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = S[r * 64 + lane];
  __syncthreads();
  wht_reg(x);  // register bits (index bits 6..11)
  // transpose: element (reg r, lane L) -> (reg L, lane r)
#pragma unroll
  for (int r = 0; r < R; ++r) S[r * SROW + lane] = x[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = S[lane * SROW + r];
  __syncthreads();
  zero_lds(S, lane);
  wht_reg(x);  // former lane bits (index bits 0..5)
}

__global__ void __launch_bounds__(64) k_compress_fwht_reg(CompressArgs a,
                                                          const uint32_t* __restrict__ EA,
                                                          const uint32_t* __restrict__ EB,
                                                          int units) {
  __shared__ __align__(16) float S[SLEN];
  const int lbs = a.lb - LB2;
  const int b1 = 1 << lbs;
  // XCD-aware numbering: workgroup g runs on XCD g % 8; give each XCD a
  // contiguous range of units so the units of one chunk share an L2
  const int G = gridDim.x;
  const int g = blockIdx.x;
  const int u = (g & 7) * (G >> 3) + (g >> 3);
  if (u >= units) return;
  const int per_chunk = a.d * b1;
  const int chunk = u / per_chunk;
  const int ts = u - chunk * per_chunk;  // t * b1 + slice
  const int t = ts >> lbs;
  const int lane = threadIdx.x;
  const int k0 = chunk * a.kper;
  const int k1 = min(a.n, k0 + a.kper);
  const float* __restrict__ AT = static_cast<const float*>(a.AT);
  const float* __restrict__ B = static_cast<const float*>(a.B);
  const uint32_t* __restrict__ ea = EA + int64_t(ts) * a.m;
  const uint32_t* __restrict__ eb = EB + int64_t(ts) * a.p;

  zero_lds(S, lane);
  __syncthreads();
  float acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.f;
  for (int k = k0; k < k1; ++k) {
    float xa[R], xb[R];
    sketch_transform(AT + int64_t(k) * a.ld_at, ea, a.m, S, lane, xa);
    __syncthreads();
    sketch_transform(B + int64_t(k) * a.ld_b, eb, a.p, S, lane, xb);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = fmaf(xa[r], xb[r], acc[r]);
  }
  (void)t;
  // lane L, register r holds frequency L*64 + r of this slice
  float4* out = reinterpret_cast<float4*>(static_cast<float*>(a.partial) +
                                          (int64_t(chunk) * per_chunk + ts) * B2 + lane * R);
#pragma unroll
  for (int r = 0; r < R; r += 4) out[r / 4] = make_float4(acc[r], acc[r + 1], acc[r + 2], acc[r + 3]);
}

}  // namespace

int compress_fwht_reg_supported(int lb) { return lb >= LB2 && lb <= 30; }

int compress_fwht_reg_lb2() { return LB2; }

size_t compress_fwht_reg_entries_bytes(int d, int lb, int m, int p) {
  return size_t(d) * (size_t(1) << (lb - LB2)) * (size_t(m) + size_t(p)) * 4;
}

hipError_t launch_compress_fwht_reg(const CompressArgs& a, void* entries, hipStream_t s) {
  if (a.nchunks == 0) return hipSuccess;
  const int lbs = a.lb - LB2;
  const int b1 = 1 << lbs;
  uint32_t* EA = static_cast<uint32_t*>(entries);
  uint32_t* EB = EA + int64_t(a.d) * b1 * a.m;
  const int64_t total = int64_t(a.d) * b1 * (int64_t(a.m) + a.p);
  if (total > 0)
    k_reg_entries<<<dim3(unsigned((total + 255) / 256)), dim3(256), 0, s>>>(a.hA, a.hB, a.m, a.p,
                                                                            a.d, lbs, EA, EB);
  const int units = a.nchunks * a.d * b1;
  const int grid = (units + 7) / 8 * 8;
  k_compress_fwht_reg<<<dim3(unsigned(grid)), dim3(64), 0, s>>>(a, EA, EB, units);
  return hipGetLastError();
}

}  // namespace cm
