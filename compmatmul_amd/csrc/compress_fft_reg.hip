// Register-resident single-wave FP32 FFT compress (the fast path for the
// cyclic-convolution variant).
//
// Same mathematics as the LDS kernel in compress.hip: per (repetition t,
// slice, chunk of k) the pair-packed sketch z = S a_k + i S' b_k of the slice
// (modulated by W_b^(h*slice)) is transformed with a 2048-point DFT, and
// Q += z^2 is accumulated (the product spectrum is recovered from Q in the
// reduce kernel).  Layout for one 64-lane wavefront:
//
//  * Element n = L + 64 r (lane L, register r < 32) after the sketch is read
//    from LDS.  Four-step DFT with n = n1 + 64 n2, k = k2 + 32 k1:
//      1. per lane, DFT-32 over the registers (DIF: one radix-2^2 step,
//         three radix-2 stages, compile-time twiddles; register q then
//         holds k2 = bitrev5(q));
//      2. twiddle by W_2048^(n1 k2) (per-lane table, L1-resident);
//      3. transpose through LDS (rows of 66 complex: 16-byte aligned, read
//         as conflict-free b128 pairs from one base address plus
//         immediates): lane L' gets k2 = L' & 31 and the 32 values
//         n1 = r' + 32 (L' >> 5);
//      4. DFT-64 over n1: one radix-2 DIF stage across lanes L' and L'+32
//         with v_permlane32_swap, a second swap that regroups the halves,
//         then a DFT-32 per lane (lanes < 32 give even k1, lanes >= 32 odd).
//  * Q (64 VGPRs) stays in registers for the chunk of k; the final sum is
//    reordered to natural frequency order through LDS.
//  * The scatter of the m + p nonzeros follows a per-(t, slice) schedule of
//    conflict-free plain LDS writes with per-item complex coefficients
//    s * W_b^(h*slice) (times i for B), built once per product.
//  * Rows k of AT and B are staged times their power-of-two balance scales
//    (balance.hip), which keep the pair packing accurate at any |a_k|/|b_k|.
#include "device_common.hpp"
#include "fft2048_reg.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

constexpr int LB2 = 11;          // slice length 2048 (complex)
constexpr int B2 = 1 << LB2;
using f2k::R;
using f2k::fq_of;
constexpr int SLEN = 2112;       // complex slots per wave: transpose (32 x 66) and dummy slots
constexpr int SEG = 1024;
constexpr int RCAP = 8;
constexpr int MAXSEG = 128;      // m + p <= MAXSEG * SEG = 131072

// ---------------------------------------------------------------- schedules
// One list per (t, slice) over the concatenated keys [0, m + p): key i < m is
// column i of A (coefficient s1(i) W_b^(h1(i) slice)), key m + j is row j of
// B (coefficient i s2(j) W_b^(h2(j) slice)).  Entry: LDS byte offset of the
// destination (bits 0..15: bucket mod 2048, or dummy slot 2048 + lane) |
// byte offset of the key in the staged segment << 16.  Per segment the groups
// are laid out round by round (round 0 plain writes, rounds 1..RCAP-2
// read-add-writes, the last round LDS atomics) with every round's group
// count in segcnt.
__host__ __device__ __forceinline__ int64_t list_cap(int keys) {
  const int64_t segs = (int64_t(keys) + SEG - 1) / SEG;
  return int64_t(keys) + segs * 64 * RCAP;
}

struct ListLayout {
  float2* C; uint32_t* L; int* segcnt; int64_t cap; int lists;
};

__host__ __device__ __forceinline__ ListLayout list_layout(const RegJob& j) {
  ListLayout o;
  o.lists = j.d << j.lbs;
  o.cap = list_cap(j.m + j.p);
  o.C = static_cast<float2*>(j.lists);
  o.L = reinterpret_cast<uint32_t*>(o.C + o.lists * o.cap);
  o.segcnt = reinterpret_cast<int*>(o.L + o.lists * o.cap);
  return o;
}

__device__ __forceinline__ void fft_lists_body(const RegJob& j, int ts) {
  const ListLayout lay = list_layout(j);
  const uint32_t* __restrict__ hA = j.hA;
  const uint32_t* __restrict__ hB = j.hB;
  const int m = j.m, p = j.p, lb = j.lb, lbs = j.lbs;
  const float2* __restrict__ tw = static_cast<const float2*>(j.tw);
  const int64_t cap = lay.cap;
  uint32_t* __restrict__ L = lay.L;
  float2* __restrict__ C = lay.C;
  int* __restrict__ segcnt = lay.segcnt;
  __shared__ uint32_t cnt[B2];
  __shared__ uint8_t rnd[SEG];
  __shared__ int nr[RCAP];
  const int b1 = 1 << lbs;
  const int t = ts >> lbs;
  const uint64_t slice = uint64_t(ts & (b1 - 1));
  const uint64_t bmask = (uint64_t(1) << lb) - 1;
  const int keys = m + p;
  uint32_t* __restrict__ Ls = L + ts * cap;
  float2* __restrict__ Cs = C + ts * cap;
  const int lane = threadIdx.x;
  for (int e = lane; e < B2; e += 64) cnt[e] = 0;
  int out = 0;
  for (int s0 = 0; s0 < keys; s0 += SEG) {
    const int sl = min(SEG, keys - s0);
    if (lane < RCAP) nr[lane] = 0;
    __syncthreads();
    for (int c = 0; c < sl; c += 64) {
      if (c + lane < sl) {
        const int key = s0 + c + lane;
        const uint32_t e = key < m ? hA[int64_t(t) * m + key] : hB[int64_t(t) * p + key - m];
        const int r = min(int(atomicAdd(&cnt[hbucket(e) & (B2 - 1)], 1u)), RCAP - 1);
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
      float2 coef = make_float2(0.f, 0.f);
      if (ok) {
        const int key = s0 + c + lane;
        const bool isb = key >= m;
        const uint32_t e = isb ? hB[int64_t(t) * p + key - m] : hA[int64_t(t) * m + key];
        const uint32_t h = hbucket(e);
        const float2 w = tw[(uint64_t(h) * slice) & bmask];
        const float s = hneg(e) ? -1.f : 1.f;
        coef = isb ? make_float2(-s * w.y, s * w.x) : make_float2(s * w.x, s * w.y);
        ent = ((h & (B2 - 1)) << 3) | (uint32_t(c + lane) << 18);
      }
#pragma unroll
      for (int rr = 0; rr < RCAP; ++rr) {
        const uint64_t mask = __ballot(r == rr);
        if (r == rr) {
          const int64_t pos = int64_t(base[rr]) * 64 + run[rr] + __popcll(mask & ((1ull << lane) - 1ull));
          Ls[pos] = ent;
          Cs[pos] = coef;
        }
        run[rr] += __popcll(mask);
      }
    }
#pragma unroll
    for (int rr = 0; rr < RCAP; ++rr) {
      const int full = (nr[rr] + 63) / 64 * 64;
      for (int slot = nr[rr] + lane; slot < full; slot += 64) {
        const int64_t pos = int64_t(base[rr]) * 64 + slot;
        Ls[pos] = uint32_t(B2 + (slot & 63)) << 3;
        Cs[pos] = make_float2(0.f, 0.f);
      }
    }
    if (lane < RCAP) {
      const int64_t row = int64_t(ts) * MAXSEG + s0 / SEG;
      segcnt[row * RCAP + lane] = (nr[lane] + 63) / 64;
    }
    out = acc;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(64) k_fft_lists(RegJob j) { fft_lists_body(j, blockIdx.x); }

__global__ void __launch_bounds__(64) k_fft_lists_batch(const RegJob* __restrict__ jobs,
                                                        const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  fft_lists_body(jobs[e.x], e.y);
}

// ---------------------------------------------------------------- transforms
// stage keys [seg*SEG, seg*SEG + SEG) of the concatenated (A row, B row),
// A's values times the balance scale bal.x and B's times bal.y
__device__ __forceinline__ void stage_rows(const float* __restrict__ ar,
                                           const float* __restrict__ br, int m, int keys,
                                           int seg, float2 bal, float* Rb, int lane) {
  const int s0 = seg * SEG;
  const int sl = min(SEG, keys - s0);
  float v[SEG / 64];
#pragma unroll
  for (int q = 0; q < SEG / 64; ++q) {
    const int key = s0 + min(q * 64 + lane, sl - 1);
    // one load from a selected address and a selected scale: a select of
    // two scaled loads compiles to a divergent branch per load
    const bool from_a = key < m;
    const float* src = from_a ? ar + key : br + (key - m);
    v[q] = *src * (from_a ? bal.x : bal.y);
  }
#pragma unroll
  for (int q = 0; q < SEG / 64; ++q) Rb[q * 64 + lane] = v[q];
}

constexpr int GB = 8;

// One round of a segment's schedule: KIND 0 plain writes, 1 read-add-
// writes, 2 LDS atomics; groups run in batches of GB whose entries,
// coefficients and staged values are loaded before the batch's LDS writes.
template <int KIND>
__device__ __forceinline__ void scatter_section(const uint32_t* __restrict__ lp,
                                                const float2* __restrict__ cp, int ng, float2* S,
                                                const float* Rb, int lane) {
  const char* Rbb = reinterpret_cast<const char*>(Rb);
  char* Sb = reinterpret_cast<char*>(S);
  for (int g0 = 0; g0 < ng; g0 += GB) {
    const int nb = min(GB, ng - g0);
    uint32_t ent[GB];
    float2 cf[GB];
    float v[GB];
#pragma unroll
    for (int q = 0; q < GB; ++q) {
      const int64_t pos = int64_t(g0 + min(q, nb - 1)) * 64 + lane;
      ent[q] = lp[pos];
      cf[q] = cp[pos];
    }
#pragma unroll
    for (int q = 0; q < GB; ++q) v[q] = *reinterpret_cast<const float*>(Rbb + ((ent[q] >> 16) & 0xffcu));
    float2 c[GB];
#pragma unroll
    for (int q = 0; q < GB; ++q) c[q] = make_float2(v[q] * cf[q].x, v[q] * cf[q].y);
#pragma unroll
    for (int q = 0; q < GB; ++q) {
      if (q < nb) {
        float2* d = reinterpret_cast<float2*>(Sb + (ent[q] & 0xffffu));
        if (KIND == 2) {
          float* df = reinterpret_cast<float*>(d);
          __hip_atomic_fetch_add(df, c[q].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(df + 1, c[q].y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else if (KIND == 1) {
          const float2 o = *d;
          *d = make_float2(o.x + c[q].x, o.y + c[q].y);
        } else {
          *d = c[q];
        }
      }
    }
  }
}

// scatter + 2048-point DFT of the pair-packed slice sketch; on return lane L,
// register q holds frequency fq_of(L, q)
__device__ __forceinline__ void sketch_fft(const float* __restrict__ ar,
                                           const float* __restrict__ br, int m, int keys,
                                           const uint32_t* __restrict__ list,
                                           const float2* __restrict__ coef,
                                           const int* __restrict__ segcnt,
                                           const float2* __restrict__ t2, float2 bal, float2* S,
                                           float* Rb, int lane, float (&re)[R], float (&im)[R]) {
  float4* S4 = reinterpret_cast<float4*>(S);
#pragma unroll
  for (int j = 0; j < B2 / 2 / 64; ++j) S4[j * 64 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
  const int nseg = (keys + SEG - 1) / SEG;
  int64_t off = 0;
  for (int seg = 0; seg < nseg; ++seg) {
    const int* __restrict__ nr = segcnt + seg * RCAP;
    __syncthreads();
    stage_rows(ar, br, m, keys, seg, bal, Rb, lane);
    __syncthreads();
    int nw = 0;
    for (int r = 1; r < RCAP - 1; ++r) nw += nr[r];
    scatter_section<0>(list + off, coef + off, nr[0], S, Rb, lane);
    off += int64_t(nr[0]) * 64;
    scatter_section<1>(list + off, coef + off, nw, S, Rb, lane);  // rounds 1..RCAP-2
    off += int64_t(nw) * 64;
    scatter_section<2>(list + off, coef + off, nr[RCAP - 1], S, Rb, lane);
    off += int64_t(nr[RCAP - 1]) * 64;
  }
  __syncthreads();
  f2k::dft2048_lds<false>(S, t2, lane, re, im);
}

__device__ __forceinline__ void fft_unit(const RegJob& j, int u, float2* S, float* Rb) {
  const ListLayout lay = list_layout(j);
  const int per_chunk = lay.lists;
  const int chunk = u / per_chunk;
  const int ts = u - chunk * per_chunk;  // t * b1 + slice
  const int lane = threadIdx.x;
  const int k0 = chunk * j.kper;
  const int k1 = min(j.n, k0 + j.kper);
  const float* __restrict__ AT = j.AT;
  const float* __restrict__ B = j.B;
  const uint32_t* __restrict__ ls = lay.L + ts * lay.cap;
  const float2* __restrict__ cs = lay.C + ts * lay.cap;
  const int* __restrict__ sc = lay.segcnt + int64_t(ts) * MAXSEG * RCAP;
  const float2* __restrict__ t2 = static_cast<const float2*>(j.t2);
  const int keys = j.m + j.p;

  float qr[R], qi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) qr[r] = qi[r] = 0.f;
  for (int k = k0; k < k1; ++k) {
    float re[R], im[R];
    sketch_fft(AT + int64_t(k) * j.ld_at, B + int64_t(k) * j.ld_b, j.m, keys, ls, cs, sc, t2,
               j.scale[k], S, Rb, lane, re, im);
#pragma unroll
    for (int r = 0; r < R; ++r) {  // Q += z^2
      qr[r] += re[r] * re[r] - im[r] * im[r];
      qi[r] = fmaf(2.f * re[r], im[r], qi[r]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) S[fq_of(lane, r)] = make_float2(qr[r], qi[r]);
  __syncthreads();
  float2* out = reinterpret_cast<float2*>(j.partial) + (int64_t(chunk) * per_chunk + ts) * B2;
#pragma unroll
  for (int r = 0; r < R; ++r) out[r * 64 + lane] = S[r * 64 + lane];
}

// Units are numbered XCD-aware (workgroup g runs on XCD g % 8 and takes a
// contiguous range of units, so the units of one chunk share an L2).  The
// loop form (one pass when the grid covers every unit) keeps the register
// allocation at two waves per SIMD.
__global__ void __launch_bounds__(64) k_compress_fft_reg(RegJob j, int units) {
  __shared__ __align__(16) float2 S[SLEN];
  __shared__ __align__(16) float Rb[SEG];
  const int G = gridDim.x;
  const int per = (units + G - 1) / G;
  const int x = blockIdx.x & 7, w = blockIdx.x >> 3;
  const int gx = G >> 3;
  for (int r = 0; r < per; ++r) {
    const int u = x * (gx * per) + r * gx + w;
    if (u >= units) break;
    fft_unit(j, u, S, Rb);
    __syncthreads();
  }
}

// Persistent form for many products (queues as in the FWHT batch kernel).
__global__ void __launch_bounds__(64) k_compress_fft_batch(const RegJob* __restrict__ jobs,
                                                           const int2* __restrict__ items,
                                                           const int* __restrict__ qstart,
                                                           unsigned* __restrict__ counters) {
  __shared__ __align__(16) float2 S[SLEN];
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
      fft_unit(jobs[e.x], e.y, S, Rb);
      __syncthreads();
    }
  }
}

}  // namespace

int fft_reg_supported(int lb, int m, int p) {
  return lb >= 1 && lb <= 30 && int64_t(m) + p <= int64_t(MAXSEG) * SEG;
}

int fft_reg_slots() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_compress_fft_reg, 64, 0) != hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return 2048;
  return cus * per_cu;
}

size_t fft_reg_lists_bytes(int d, int lb, int m, int p) {
  const size_t lists = size_t(d) << reg_lbs(lb);
  return lists * size_t(list_cap(m + p)) * 12 + lists * MAXSEG * RCAP * sizeof(int) + 256;
}

hipError_t launch_fft_reg(const RegJob& j, hipStream_t s) {
  if (j.nchunks == 0) return hipSuccess;
  const int lists = j.d << j.lbs;
  k_fft_lists<<<dim3(unsigned(lists)), dim3(64), 0, s>>>(j);
  const int units = j.nchunks * lists;
  const int grid = (units + 7) / 8 * 8;
  k_compress_fft_reg<<<dim3(unsigned(grid)), dim3(64), 0, s>>>(j, units);
  return hipGetLastError();
}

hipError_t launch_fft_reg_lists_batch(const RegJob* jobs, const int2* blocks, int nblocks,
                                      hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  k_fft_lists_batch<<<dim3(unsigned(nblocks)), dim3(64), 0, s>>>(jobs, blocks);
  return hipGetLastError();
}

hipError_t launch_fft_reg_batch(const RegJob* jobs, const int2* items, const int* qstart,
                                unsigned* counters, int grid, hipStream_t s) {
  if (grid <= 0) return hipSuccess;
  k_compress_fft_batch<<<dim3(unsigned(grid)), dim3(64), 0, s>>>(jobs, items, qstart, counters);
  return hipGetLastError();
}

}  // namespace cm
