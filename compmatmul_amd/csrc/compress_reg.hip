// Register-resident single-wave FP32 FWHT compress (the fast path).
//
// Same mathematics as the LDS kernel in compress.hip (sliced, modulated
// count-sketch into b2 = 2048 buckets, natural-order Walsh-Hadamard
// transform, pointwise product, accumulation over a chunk of k), laid out
// for one 64-lane wavefront per workgroup:
//
//  * A slice of 2048 values lives in 32 VGPRs per lane.  The WHT is a tensor
//    product of 2-point butterflies over the 11 index bits: after the sketch
//    is read from LDS (lane L, register r = element r*64 + L) the five
//    register bits are transformed in registers; a transpose through LDS
//    (rows of 33 floats: conflict-free, one base address plus immediates)
//    moves index bits 0..4 into registers and bit 5 into lane bit 5; those
//    five are transformed in registers and the last bit with
//    v_permlane32_swap (lanes L and L+32) and one butterfly.
//  * The spectral accumulator (32 VGPRs) stays in registers for the whole
//    chunk of k; only the final sum is written (reordered through LDS).
//  * The scatter follows a per-(t, slice, operand) schedule of
//    conflict-free plain LDS writes built once per product (k_reg_lists):
//    LDS float atomics cost ~190 clk per wave instruction on gfx950, plain
//    random LDS writes ~8.  The row is staged in LDS 1024 keys at a time, so
//    each scheduled item is an LDS gather plus a write.
//  * LDS per wave: 8.25 KB sketch/transpose + 4 KB row stage; about 160
//    VGPRs: three waves per SIMD.
//  * Workgroups are numbered so that the (t, slice) units of one chunk of k
//    run on the same XCD and share its L2 copy of the rows of AT and B.
#include "device_common.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

constexpr int LB2 = 11;           // slice length 2048
constexpr int B2 = 1 << LB2;
constexpr int R = B2 / 64;        // values per lane (32)
constexpr int TROW = 33;          // padded row (32 floats) of the transpose
constexpr int SLEN = 64 * TROW;   // LDS floats per wave for sketch/transpose

__device__ __forceinline__ float flip(float v, uint32_t sgn) {
  return __int_as_float(__float_as_int(v) ^ sgn);
}

// ---------------------------------------------------------------- schedules
// LDS float atomics are slow on gfx950 (~190 clk per wave instruction), so
// the scatter is a precomputed schedule of plain LDS writes.  For each
// (t, slice, operand) the keys of a row are processed in segments of 1024;
// within a segment, items are grouped by their occurrence index in their
// bucket ("round"): round 0 (first occurrence in the whole row) writes its
// value, later rounds read-add-write, and ranks >= RCAP-1 fall back to LDS
// atomics.  Items of one group (64 lanes) never share a bucket, and every
// group runs after the groups holding earlier items of its buckets, so the
// result is exact FP32 in a fixed order (deterministic).
// Entry: LDS byte offset of the destination (bits 0..15: bucket mod 2048,
// or a per-lane dummy slot 2048+lane for padding) | byte offset of the key
// in the staged segment << 16 (bits 16..27) | sign << 31.  Per segment the
// groups are laid out round by round (round 0 plain writes, rounds
// 1..RCAP-2 read-add-writes, the last round LDS atomics) and the group
// count of every round is stored.
constexpr int SEG = 1024;
constexpr int RCAP = 8;
constexpr int MAXSEG = 64;  // keys <= MAXSEG * SEG = 65536
constexpr uint32_t F_SIGN = 1u << 31;

__host__ __device__ __forceinline__ int64_t list_cap(int keys) {
  const int64_t segs = (int64_t(keys) + SEG - 1) / SEG;
  return int64_t(keys) + segs * 64 * RCAP;
}

struct ListLayout {
  uint32_t* LA; uint32_t* LB; int* ngroups; int64_t cap_a, cap_b; int lists;
};

__host__ __device__ __forceinline__ ListLayout list_layout(const RegJob& j) {
  ListLayout o;
  o.lists = j.d << j.lbs;
  o.cap_a = list_cap(j.m);
  o.cap_b = list_cap(j.p);
  o.LA = static_cast<uint32_t*>(j.lists);
  o.LB = o.LA + o.lists * o.cap_a;
  o.ngroups = reinterpret_cast<int*>(o.LB + o.lists * o.cap_b);
  return o;
}

// schedule of list g: g < lists covers A's row for (t, slice) = g, the rest B's
__device__ __forceinline__ void reg_lists_body(const RegJob& j, int g) {
  const ListLayout lay = list_layout(j);
  const int lbs = j.lbs, m = j.m, p = j.p;
  const uint32_t* __restrict__ hA = j.hA;
  const uint32_t* __restrict__ hB = j.hB;
  const int64_t cap_a = lay.cap_a, cap_b = lay.cap_b;
  uint32_t* __restrict__ LA = lay.LA;
  uint32_t* __restrict__ LB = lay.LB;
  int* __restrict__ ngroups = lay.ngroups;
  __shared__ uint32_t cnt[B2];
  __shared__ uint8_t rnd[SEG];
  __shared__ int nr[RCAP];
  const int b1 = 1 << lbs;
  const int lists = lay.lists;
  const bool side_b = g >= lists;
  const int ts = side_b ? g - lists : g;
  const int t = ts >> lbs;
  const uint32_t slice = uint32_t(ts & (b1 - 1));
  const int keys = side_b ? p : m;
  const uint32_t* __restrict__ hh = (side_b ? hB : hA) + int64_t(t) * keys;
  uint32_t* __restrict__ L = side_b ? LB + ts * cap_b : LA + ts * cap_a;
  const int lane = threadIdx.x;
  for (int e = lane; e < B2; e += 64) cnt[e] = 0;
  int out = 0;  // groups written so far
  for (int s0 = 0; s0 < keys; s0 += SEG) {
    const int sl = min(SEG, keys - s0);
    if (lane < RCAP) nr[lane] = 0;
    __syncthreads();
    for (int c = 0; c < sl; c += 64) {
      if (c + lane < sl) {
        const uint32_t h = hbucket(hh[s0 + c + lane]);
        const int r = min(int(atomicAdd(&cnt[h & (B2 - 1)], 1u)), RCAP - 1);
        rnd[c + lane] = uint8_t(r);
        atomicAdd(&nr[r], 1);
      }
      __syncthreads();
    }
    int base[RCAP], run[RCAP];
    int acc = out;
#pragma unroll
    for (int r = 0; r < RCAP; ++r) {
      base[r] = acc;
      acc += (nr[r] + 63) / 64;
      run[r] = 0;
    }
    for (int c = 0; c < sl; c += 64) {
      const bool ok = c + lane < sl;
      const int r = ok ? int(rnd[c + lane]) : -1;
      uint32_t ent = 0;
      if (ok) {
        const uint32_t e = hh[s0 + c + lane];
        const uint32_t h = hbucket(e);
        const uint32_t neg = hneg(e) ^ (__popc((h >> LB2) & slice) & 1u);
        ent = ((h & (B2 - 1)) << 2) | (uint32_t(c + lane) << 18) | (neg ? F_SIGN : 0u);
      }
#pragma unroll
      for (int rr = 0; rr < RCAP; ++rr) {
        const uint64_t mask = __ballot(r == rr);
        if (r == rr) {
          const int pos = run[rr] + __popcll(mask & ((1ull << lane) - 1ull));
          L[int64_t(base[rr]) * 64 + pos] = ent;
        }
        run[rr] += __popcll(mask);
      }
    }
    // pad every round to whole groups with writes to per-lane dummy slots
#pragma unroll
    for (int rr = 0; rr < RCAP; ++rr) {
      const int full = (nr[rr] + 63) / 64 * 64;
      for (int slot = nr[rr] + lane; slot < full; slot += 64)
        L[int64_t(base[rr]) * 64 + slot] =
            uint32_t(B2 + (slot & 63)) << 2;
    }
    if (lane < RCAP) {
      const int64_t row = int64_t(g) * MAXSEG + s0 / SEG;
      ngroups[row * RCAP + lane] = (nr[lane] + 63) / 64;
    }
    out = acc;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(64) k_reg_lists(RegJob j) { reg_lists_body(j, blockIdx.x); }

__global__ void __launch_bounds__(64) k_reg_lists_batch(const RegJob* __restrict__ jobs,
                                                        const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  reg_lists_body(jobs[e.x], e.y);
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

constexpr int GB = 16;  // schedule groups per batch

// stage keys [seg*SEG, seg*SEG + SEG) of a row into Rb
__device__ __forceinline__ void stage_row(const float* __restrict__ row, int keys, int seg,
                                          float* Rb, int lane) {
  const int s0 = seg * SEG;
  const int sl = min(SEG, keys - s0);
  float v[SEG / 64];
#pragma unroll
  for (int q = 0; q < SEG / 64; ++q) v[q] = row[s0 + min(q * 64 + lane, sl - 1)];
#pragma unroll
  for (int q = 0; q < SEG / 64; ++q) Rb[q * 64 + lane] = v[q];
}

// One round of a segment's schedule: KIND 0 plain writes, 1 read-add-
// writes, 2 LDS atomics; groups run in batches of GB whose entries and
// staged values are loaded before the batch's LDS writes.
template <int KIND>
__device__ __forceinline__ void scatter_section(const uint32_t* __restrict__ lp, int ng, float* S,
                                                const float* Rb, int lane) {
  const char* Rbb = reinterpret_cast<const char*>(Rb);
  char* Sb = reinterpret_cast<char*>(S);
  for (int g0 = 0; g0 < ng; g0 += GB) {
    const int nb = min(GB, ng - g0);
    uint32_t ent[GB];
    float v[GB];
#pragma unroll
    for (int q = 0; q < GB; ++q) ent[q] = lp[int64_t(g0 + min(q, nb - 1)) * 64 + lane];
#pragma unroll
    for (int q = 0; q < GB; ++q)
      v[q] = flip(*reinterpret_cast<const float*>(Rbb + ((ent[q] >> 16) & 0xffcu)), ent[q] & F_SIGN);
#pragma unroll
    for (int q = 0; q < GB; ++q) {
      if (q < nb) {
        float* d = reinterpret_cast<float*>(Sb + (ent[q] & 0xffffu));
        if (KIND == 2)
          __hip_atomic_fetch_add(d, v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else if (KIND == 1)
          *d += v[q];
        else
          *d = v[q];
      }
    }
  }
}

// scatter one row into S following its schedule, read the sketch into x,
// full WHT of x (the transpose leaves S dirty; the next scatter zeroes it).
// On return lane L, register r holds frequency f_of(L, r).
__device__ __forceinline__ void sketch_transform(const float* __restrict__ row, int keys,
                                                 const uint32_t* __restrict__ list,
                                                 const int* __restrict__ segcnt, float* S,
                                                 float* Rb, int lane, float (&x)[R]) {
  float4* S4 = reinterpret_cast<float4*>(S);
#pragma unroll
  for (int j = 0; j < B2 / 4 / 64; ++j) S4[j * 64 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
  const int nseg = (keys + SEG - 1) / SEG;
  const uint32_t* __restrict__ lp = list;
  for (int seg = 0; seg < nseg; ++seg) {
    const int* __restrict__ nr = segcnt + seg * RCAP;
    __syncthreads();
    stage_row(row, keys, seg, Rb, lane);
    __syncthreads();
    int nw = 0;
    for (int r = 1; r < RCAP - 1; ++r) nw += nr[r];
    scatter_section<0>(lp, nr[0], S, Rb, lane);
    lp += int64_t(nr[0]) * 64;
    scatter_section<1>(lp, nw, S, Rb, lane);  // rounds 1..RCAP-2, in order
    lp += int64_t(nw) * 64;
    scatter_section<2>(lp, nr[RCAP - 1], S, Rb, lane);
    lp += int64_t(nr[RCAP - 1]) * 64;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = S[r * 64 + lane];
  __syncthreads();
  wht_reg(x);  // register bits: index bits 6..10
  // transpose: element e = r*64 + L goes to row e >> 5, column e & 31; lane
  // L then reads row ((L & 31) << 1) | (L >> 5), i.e. index bits 0..4 land in
  // registers, bits 6..10 in lane bits 0..4 and bit 5 in lane bit 5
  {
    float* wbase = S + (lane >> 5) * TROW + (lane & 31);
#pragma unroll
    for (int r = 0; r < R; ++r) wbase[r * 2 * TROW] = x[r];
  }
  __syncthreads();
  {
    const float* rbase = S + (((lane & 31) << 1) | (lane >> 5)) * TROW;
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = rbase[r];
  }
  __syncthreads();
  wht_reg(x);  // index bits 0..4
  // index bit 5 sits in lane bit 5: swap halves between register pairs and
  // butterfly (lanes < 32 finish register 2j's pairs, lanes >= 32 2j+1's)
#pragma unroll
  for (int j = 0; j < R / 2; ++j) {
    const auto pq = __builtin_amdgcn_permlane32_swap(__float_as_uint(x[2 * j]),
                                                     __float_as_uint(x[2 * j + 1]), false, false);
    const float p = __uint_as_float(pq[0]), q = __uint_as_float(pq[1]);
    x[2 * j] = p + q;
    x[2 * j + 1] = p - q;
  }
}

// frequency held by lane L, register r after sketch_transform
__device__ __forceinline__ int f_of(int L, int r) {
  return ((L & 31) << 6) | ((r & 1) << 5) | ((r & ~1) + (L >> 5));
}

// One unit: (chunk of k, t, slice) = u of job j, over k in [chunk * kper, ...).
__device__ __forceinline__ void fwht_unit(const RegJob& j, int u, float* S, float* Rb) {
  const ListLayout lay = list_layout(j);
  const int per_chunk = lay.lists;
  const int chunk = u / per_chunk;
  const int ts = u - chunk * per_chunk;  // t * b1 + slice
  const int lane = threadIdx.x;
  const int k0 = chunk * j.kper;
  const int k1 = min(j.n, k0 + j.kper);
  const float* __restrict__ AT = j.AT;
  const float* __restrict__ B = j.B;
  const uint32_t* __restrict__ la = lay.LA + ts * lay.cap_a;
  const uint32_t* __restrict__ lb = lay.LB + ts * lay.cap_b;
  const int* __restrict__ sca = lay.ngroups + int64_t(ts) * MAXSEG * RCAP;
  const int* __restrict__ scb = lay.ngroups + int64_t(per_chunk + ts) * MAXSEG * RCAP;

  float acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.f;
  for (int k = k0; k < k1; ++k) {
    float xa[R], xb[R];
    sketch_transform(AT + int64_t(k) * j.ld_at, j.m, la, sca, S, Rb, lane, xa);
    sketch_transform(B + int64_t(k) * j.ld_b, j.p, lb, scb, S, Rb, lane, xb);
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = fmaf(xa[r], xb[r], acc[r]);
  }
  // reorder through LDS to natural frequency order, then coalesced stores
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) S[f_of(lane, r)] = acc[r];
  __syncthreads();
  float* out = j.partial + (int64_t(chunk) * per_chunk + ts) * B2;
#pragma unroll
  for (int r = 0; r < R; ++r) out[r * 64 + lane] = S[r * 64 + lane];
}

__global__ void __launch_bounds__(64) k_compress_fwht_reg(RegJob j, int units) {
  __shared__ __align__(16) float S[SLEN];
  __shared__ __align__(16) float Rb[SEG];
  // XCD-aware numbering: workgroup g runs on XCD g % 8; give each XCD a
  // contiguous range of units so the units of one chunk share an L2
  const int G = gridDim.x;
  const int g = blockIdx.x;
  const int u = (g & 7) * (G >> 3) + (g >> 3);
  if (u >= units) return;
  fwht_unit(j, u, S, Rb);
}

// Persistent form for many products: every workgroup takes items from the
// queue of its XCD until it is empty, then from the other queues in turn.
__global__ void __launch_bounds__(64) k_compress_fwht_batch(const RegJob* __restrict__ jobs,
                                                            const int2* __restrict__ items,
                                                            const int* __restrict__ qstart,
                                                            unsigned* __restrict__ counters) {
  __shared__ __align__(16) float S[SLEN];
  __shared__ __align__(16) float Rb[SEG];
  const int home = blockIdx.x & 7;
  for (int qi = 0; qi < 8; ++qi) {
    const int q = (home + qi) & 7;
    const int base = qstart[q];
    const int len = qstart[q + 1] - base;
    int last = -1;
    for (;;) {
      unsigned it = 0;
      if (threadIdx.x == 0) it = claim_item(&counters[q], len, probe_next(qi, len, last));
      it = __builtin_amdgcn_readfirstlane(it);
      if (int(it) >= len) break;
      last = int(it);
      const int2 e = items[base + int(it)];
      fwht_unit(jobs[e.x], e.y, S, Rb);
      __syncthreads();  // the next unit rewrites S and Rb
    }
  }
}

__global__ void __launch_bounds__(256) k_zero_u32(unsigned* __restrict__ p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace

int fwht_reg_supported(int lb, int m, int p) {
  return lb >= 1 && lb <= 30 && m <= MAXSEG * SEG && p <= MAXSEG * SEG;
}

int fwht_reg_slots() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_compress_fwht_reg, 64, 0) != hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return 2048;
  return cus * per_cu;
}

size_t fwht_reg_lists_bytes(int d, int lb, int m, int p) {
  const size_t lists = size_t(d) << reg_lbs(lb);
  return lists * size_t(list_cap(m) + list_cap(p)) * 4 + 2 * lists * MAXSEG * RCAP * sizeof(int) + 256;
}

hipError_t launch_fwht_reg(const RegJob& j, hipStream_t s) {
  if (j.nchunks == 0) return hipSuccess;
  const int lists = j.d << j.lbs;
  k_reg_lists<<<dim3(unsigned(2 * lists)), dim3(64), 0, s>>>(j);
  const int units = j.nchunks * lists;
  const int grid = (units + 7) / 8 * 8;
  k_compress_fwht_reg<<<dim3(unsigned(grid)), dim3(64), 0, s>>>(j, units);
  return hipGetLastError();
}

hipError_t launch_fwht_reg_lists_batch(const RegJob* jobs, const int2* blocks, int nblocks,
                                       hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  k_reg_lists_batch<<<dim3(unsigned(nblocks)), dim3(64), 0, s>>>(jobs, blocks);
  return hipGetLastError();
}

hipError_t launch_fwht_reg_batch(const RegJob* jobs, const int2* items, const int* qstart,
                                 unsigned* counters, int grid, hipStream_t s) {
  if (grid <= 0) return hipSuccess;
  k_compress_fwht_batch<<<dim3(unsigned(grid)), dim3(64), 0, s>>>(jobs, items, qstart, counters);
  return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_zero_u32<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, s>>>(p, n);
  return hipGetLastError();
}

}  // namespace cm
