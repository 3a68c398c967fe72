// Median-of-d decompression and large-entry recovery.
//
// Reference: matmul_decompress_fft (matmul_fft.hpp:135-147) and
// matmul_decompress_fwht (matmul_fwht.hpp:122-134) evaluated for every
// (i, j) (matmul_fft.hpp:165-177), with array::median = nth_element at d/2
// (array.hpp:323-328; the upper median for even d).
//   Chat[i,j] = median_t( s1_t(i) s2_t(j) p[t, h] ),
//   h = (h1_t(i) + h2_t(j)) & (b-1)  (FFT)   or   h1_t(i) ^ h2_t(j)  (FWHT).
// One thread owns a column j for a group of rows: the d packed hash entries
// of column j stay in registers, the row entries are wave-uniform loads,
// the sketch (d x b) is gathered from L2, and the median comes from a
// register selection network (d <= 16) or a rank count (larger d).
// Recovery appends (i, j, value) for every estimate > tau through a global
// counter (one atomic per wave in the staged kernels); the host sorts the
// list by (i, j).
#include <type_traits>

#include <atomic>

#include "device_common.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

constexpr int NT = 256;
constexpr int ROWS = 8;

// f(integral_constant<int, I>) for I = I0 .. N-1: a loop whose index is a
// constant in every copy of the body, however large the body grows (an
// unroll pragma gives up past a size threshold, and then arrays indexed by
// the loop variable leave registers for scratch memory)
template <int I0, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I0 < N) {
    f(std::integral_constant<int, I0>{});
    static_for<I0 + 1, N>(f);
  }
}

__device__ __forceinline__ float vmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float vmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double vmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double vmax(double a, double b) { return fmax(a, b); }

// Median selection network for D <= 16 values, built at compile time:
// Batcher's odd-even merge sort of the next power of two N (the padding
// positions D..N-1 act as +inf, so comparators touching them are dropped),
// then pruned backwards to the comparators that can move a value into
// output D/2.  Compare-exchanges only move values, so v[D/2] afterwards is
// exactly element D/2 of the sorted values.  D = 15: 49 comparators (a full
// odd-even transposition sort takes 105).
struct CmpNet {
  int n;
  unsigned char a[64], b[64];
};

// Batcher's odd-even merge sort of D values (comparator c orders a[c] before
// b[c]), padded to the next power of two as above.
constexpr CmpNet sort_net(int D) {
  CmpNet full{0, {}, {}};
  int N = 1;
  while (N < D) N *= 2;
  for (int p = 1; p < N; p *= 2)
    for (int k = p; k >= 1; k /= 2)
      for (int j = k % p; j + k < N; j += 2 * k)
        for (int i = 0; i < (k < N - j - k ? k : N - j - k); ++i)
          if ((i + j) / (2 * p) == (i + j + k) / (2 * p) && i + j + k < D) {
            full.a[full.n] = static_cast<unsigned char>(i + j);
            full.b[full.n] = static_cast<unsigned char>(i + j + k);
            ++full.n;
          }
  return full;
}

constexpr CmpNet median_net(int D) {
  const CmpNet full = sort_net(D);
  bool need[16] = {};
  bool keep[64] = {};
  need[D / 2] = true;
  for (int c = full.n - 1; c >= 0; --c)
    if (need[full.a[c]] || need[full.b[c]]) {
      keep[c] = true;
      need[full.a[c]] = need[full.b[c]] = true;
    }
  CmpNet out{0, {}, {}};
  for (int c = 0; c < full.n; ++c)
    if (keep[c]) {
      out.a[out.n] = full.a[c];
      out.b[out.n] = full.b[c];
      ++out.n;
    }
  return out;
}

template <typename T, int D>
__device__ __forceinline__ T median_network(T (&v)[D]) {
  static_assert(D >= 1 && D <= 16, "selection network for d <= 16");
  constexpr CmpNet net = median_net(D);
#pragma unroll
  for (int c = 0; c < net.n; ++c) {
    const int i = net.a[c], k = net.b[c];
    const T lo = vmin(v[i], v[k]);
    const T hi = vmax(v[i], v[k]);
    v[i] = lo;
    v[k] = hi;
  }
  return v[D / 2];
}

// x clamped to [lo, hi] (lo <= hi): the middle one of the three values,
// returned unchanged (one v_med3_f32)
__device__ __forceinline__ float vclamp(float x, float lo, float hi) {
  return __builtin_amdgcn_fmed3f(x, lo, hi);
}

// Streaming form of the same selection, for kernels that receive the D values
// of an entry one repetition at a time.  Element D/2 of the ascending order
// is the K-th largest value, K = D - D/2, so only the K largest values seen
// so far are kept, in descending order in a[0..K-1]:
// - values 0..K-1 are stored, then sorted with Batcher's network of K;
// - value t >= K, the q-th of the R = D - K later values, enters as
//   a[k] = clamp(x, a[k], a[k-1]) for k = K-1 down to 1, and a[0] = max(a[0], x).
//   A value moves at most one place per later insertion, so only positions
//   k >= K-1-(R-q) can still reach a[K-1]; the others are left stale and
//   never read again.
// Every step moves values without arithmetic, so a[K-1] at the end is
// exactly element D/2 of the sorted values.  D = 15: 19 comparators and 28
// clamps (the pruned network takes 49 comparators).  The selection work then
// runs between the row copies instead of after the last one.  Used in FP32
// for D >= 10, where it needs no more registers than the network (FP64 has
// no med3 instruction, and at smaller D the compiler's schedule for it costs
// occupancy or spills).
template <typename T, int D>
constexpr bool streamed() {
  return sizeof(T) == 4 && D >= 10;
}

// values kept per entry: the K largest when streamed, else all D
template <typename T, int D>
constexpr int keep_count() {
  return streamed<T, D>() ? D - D / 2 : D;
}

// value t of an entry into its kept values a
template <typename T, int D, int t, int N>
__device__ __forceinline__ void select_push(T (&a)[N], T x) {
  static_assert(N == keep_count<T, D>(), "kept values per entry");
  if constexpr (!streamed<T, D>()) {
    a[t] = x;
    return;
  }
  constexpr int K = keep_count<T, D>(), R = D - K;
  if constexpr (t < K) {
    a[t] = x;
    if constexpr (t == K - 1) {
      constexpr CmpNet net = sort_net(K);
#pragma unroll
      for (int c = 0; c < net.n; ++c) {
        const int i = net.a[c], k = net.b[c];
        const T hi = vmax(a[i], a[k]);
        const T lo = vmin(a[i], a[k]);
        a[i] = hi;
        a[k] = lo;
      }
    }
  } else {
    constexpr int q = t - K + 1;
    constexpr int lo = K - 1 - (R - q) > 0 ? K - 1 - (R - q) : 0;
#pragma unroll
    for (int k = K - 1; k >= (lo > 1 ? lo : 1); --k) a[k] = vclamp(x, a[k], a[k - 1]);
    if constexpr (lo == 0) a[0] = vmax(a[0], x);
  }
}

// element D/2 of an entry's sorted values, from its kept values
template <typename T, int D, int N>
__device__ __forceinline__ T select_result(T (&a)[N]) {
  static_assert(N == keep_count<T, D>(), "kept values per entry");
  if constexpr (streamed<T, D>())
    return a[keep_count<T, D>() - 1];
  else
    return median_network<T, D>(a);
}

template <typename T, bool FWHT>
__device__ __forceinline__ T gather(const T* __restrict__ P, uint32_t ea, uint32_t eb,
                                    uint32_t mask, int64_t toff) {
  const uint32_t h = FWHT ? (hbucket(ea) ^ hbucket(eb)) : ((hbucket(ea) + hbucket(eb)) & mask);
  const T v = P[toff + h];
  return (hneg(ea) ^ hneg(eb)) ? -v : v;
}

template <typename T>
__device__ __forceinline__ void emit(const DecompressArgs& a, int i, int j, T med) {
  if (a.C) static_cast<T*>(a.C)[int64_t(i) * a.ldc + j] = med;
  if (a.recover && double(med) > a.tau) {
    const unsigned long long slot = atomicAdd(a.counter, 1ull);
    if (int64_t(slot) < a.max_out) {
      a.oi[slot] = i;
      a.oj[slot] = j;
      static_cast<T*>(a.ov)[slot] = med;
    }
  }
}

// The staged kernels' form of emit for a thread's ER rows of column j: C is
// written, and the estimates above tau are appended with one atomic per
// wave.  Each lane counts its passing rows, a prefix sum over the wave
// places them, and lane 63 claims the wave's slots.  Every lane of the wave
// calls it (rows outside C have `jok` false or i >= m).  An atomic per
// (wave, row) serialized on the counter when most estimates pass (small b
// and large m p): about 40 ns each, milliseconds per product.
template <typename T, int ER>
__device__ __forceinline__ void emit_rows(const DecompressArgs& a, int i0, int j, bool jok,
                                          const T (&med)[ER]) {
  unsigned pass = 0;
#pragma unroll
  for (int e = 0; e < ER; ++e) {
    const bool in = jok && i0 + e < a.m;
    if (in && a.C) static_cast<T*>(a.C)[int64_t(i0 + e) * a.ldc + j] = med[e];
    if (in && a.recover && double(med[e]) > a.tau) pass |= 1u << e;
  }
  if (!a.recover) return;
  const int lane = int(threadIdx.x) & 63;
  const int cnt = __popc(pass);
  int incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  const int total = __shfl(incl, 63, 64);
  if (total == 0) return;
  unsigned long long base = 0;
  if (lane == 63) base = atomicAdd(a.counter, (unsigned long long)total);
  unsigned long long slot = __shfl(base, 63, 64) + unsigned(incl - cnt);
#pragma unroll
  for (int e = 0; e < ER; ++e) {
    if (!((pass >> e) & 1u)) continue;
    if (int64_t(slot) < a.max_out) {
      a.oi[slot] = i0 + e;
      a.oj[slot] = j;
      static_cast<T*>(a.ov)[slot] = med[e];
    }
    ++slot;
  }
}

template <typename T, bool FWHT, int D>
__device__ __forceinline__ void decompress_body(const DecompressArgs& a, int blk) {
  const int tiles = (a.p + NT - 1) / NT;
  const int j = int(blk % tiles) * NT + threadIdx.x;
  const int i0 = int(blk / tiles) * ROWS;
  if (j >= a.p) return;
  const T* __restrict__ P = static_cast<const T*>(a.P);
  const uint32_t mask = uint32_t(a.b - 1);
  uint32_t eb[D];
#pragma unroll
  for (int t = 0; t < D; ++t) eb[t] = a.hB[int64_t(t) * a.p + j];
  const int i1 = min(a.m, i0 + ROWS);
  for (int i = i0; i < i1; ++i) {
    T v[D];
#pragma unroll
    for (int t = 0; t < D; ++t)
      v[t] = gather<T, FWHT>(P, a.hA[int64_t(t) * a.m + i], eb[t], mask, int64_t(t) * a.b);
    emit<T>(a, i, j, median_network<T, D>(v));
  }
}

template <typename T, bool FWHT, int D>
__global__ void __launch_bounds__(NT) k_decompress(DecompressArgs a) {
  decompress_body<T, FWHT, D>(a, blockIdx.x);
}

// LDS-staged variant for sketch rows of at most 64 KB: the workgroup (512
// threads, one column each, up to 12 rows per thread) streams the d rows of
// the sketch through two LDS buffers, so the random gathers hit LDS instead
// of 128-byte L2 lines.  Rows are copied with global_load_lds (straight into
// LDS, 1 KB per wave instruction, no registers): row t+1 is in flight while
// row t is gathered, and the registers go to rows of C instead.
constexpr int LNT = 512;
// rows per thread: about 120 gathered values per thread
template <typename T, int D>
constexpr int rows_per_thread() {
  return (sizeof(T) == 8 ? 60 : 120) / D < 1 ? 1 : ((sizeof(T) == 8 ? 60 : 120) / D > 12 ? 12 : (sizeof(T) == 8 ? 60 : 120) / D);
}

// start copying sketch row t into LDS buffer t & 1
template <typename T>
__device__ __forceinline__ void stage_row_async(const T* __restrict__ P, T* buf0, int b, int t,
                                                int tid) {
  const int vec = b * int(sizeof(T)) / 16;  // float4 per row
  const int wave = tid >> 6, lane = tid & 63;
  const float4* src = reinterpret_cast<const float4*>(P + int64_t(t) * b);
  float4* dst = reinterpret_cast<float4*>(buf0 + (t & 1) * b);
  for (int q0 = wave * 64; q0 < vec; q0 += LNT) {
    if (q0 + lane < vec)  // lanes past the row write nothing
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + q0 + lane),
                                       (__attribute__((address_space(3))) void*)(dst + q0), 16, 0, 0);
  }
}

// val negated when the sign bits of entries ea and eb differ, as one xor of
// the IEEE sign bit (exactly what negation does), so the signed value is
// formed at once rather than the entries being kept live until its use.
__device__ __forceinline__ float signed_by(float val, uint32_t ea, uint32_t eb) {
  return __int_as_float(__float_as_int(val) ^ int((ea ^ eb) & 0x80000000u));
}
__device__ __forceinline__ double signed_by(double val, uint32_t ea, uint32_t eb) {
  return __hiloint2double(__double2hiint(val) ^ int((ea ^ eb) & 0x80000000u), __double2loint(val));
}

// Hash entry i of a read-only table, through the constant address space:
// at a uniform index this is a scalar load, counted by lgkmcnt, so the waits
// for it never wait for the row copies in flight (vmcnt), as a vector load
// issued after them would.
__device__ __forceinline__ uint32_t hash_entry(const uint32_t* h, int64_t i) {
  return ((const __attribute__((address_space(4))) uint32_t*)h)[i];
}

template <typename T, bool FWHT, int D>
__device__ __forceinline__ void decompress_lds_body(const DecompressArgs& a, int blk) {
  static_assert(LNT % 64 == 0, "emit_rows sums over whole waves");
  constexpr int ER = rows_per_thread<T, D>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = a.b;
  T* const buf0 = reinterpret_cast<T*>(smem_raw);
  const int tiles = (a.p + LNT - 1) / LNT;
  const int tid = threadIdx.x;
  const int j = int(blk % tiles) * LNT + tid;
  const int i0 = int(blk / tiles) * ER;
  const bool jok = j < a.p;
  const uint32_t mask = uint32_t(b - 1);
  const T* __restrict__ P = static_cast<const T*>(a.P);
  uint32_t eb[D];
#pragma unroll
  for (int t = 0; t < D; ++t) eb[t] = jok ? a.hB[int64_t(t) * a.p + j] : 0u;
  stage_row_async<T>(P, buf0, b, 0, tid);
  T v[ER][keep_count<T, D>()];
  static_for<0, D>([&](auto tc) {
    constexpr int t = decltype(tc)::value;
    __builtin_amdgcn_s_waitcnt(0);  // this thread's copies of row t are done
    __syncthreads();                // everyone's, and row t-1 is no longer read
    if (t + 1 < D) stage_row_async<T>(P, buf0, b, t + 1, tid);
    const T* cur = buf0 + (t & 1) * b;
    const uint32_t* hA = a.hA;
    asm volatile("" : "+s"(hA));  // row t's entries are loaded here, not all hoisted to the start
    T x[ER];
#pragma unroll
    for (int e = 0; e < ER; ++e) {
      const int i = min(i0 + e, a.m - 1);
      const uint32_t ea = hash_entry(hA, int64_t(t) * a.m + i);
      const uint32_t h = FWHT ? (hbucket(ea) ^ hbucket(eb[t])) : ((hbucket(ea) + hbucket(eb[t])) & mask);
      const T val = cur[h];
      x[e] = signed_by(val, ea, eb[t]);
    }
    // formed here, after all of the row's gathers are issued: the entries
    // are not kept live until the selection
#pragma unroll
    for (int e = 0; e < ER; ++e) asm volatile("" : "+v"(x[e]));
#pragma unroll
    for (int e = 0; e < ER; ++e) select_push<T, D, t>(v[e], x[e]);
  });
  T med[ER];
#pragma unroll
  for (int e = 0; e < ER; ++e) med[e] = select_result<T, D>(v[e]);
  emit_rows<T, ER>(a, i0, j, jok, med);
}

// Ring variant for 64 KB and 32 KB sketch rows (the two-buffer form above
// keeps one row in flight): five 32 KB slots fill the 160 KB of LDS.  A
// 64 KB row is two halves, half k of the d x 2 goes to slot k mod 5, and
// while row t is gathered the halves 2t+2 .. 2t+4 are in flight (1.5 rows);
// a 32 KB row takes one slot, and rows t+1 .. t+4 are in flight.  Row t's
// slots are refilled after the barrier of row t + 1, so one barrier per row
// still suffices.
constexpr int RING_SLOTS = 5;
constexpr int RING_HALF = 32768;                  // bytes per half row
constexpr int RING_LDS = RING_SLOTS * RING_HALF;  // 160 KB
// dwordx4 copies per wave and half: 4 with 512 threads, 2 with 1024
template <int NT>
constexpr int ring_issue() {
  return RING_HALF / 16 / NT;
}

// copy 32 KB piece k (piece k % HP of row k / HP) into slot k mod 5
template <typename T, int NT, int HP>
__device__ __forceinline__ void stage_half_async(const T* __restrict__ P, unsigned char* ring, int b,
                                                 int k, int tid) {
  const unsigned char* src =
      reinterpret_cast<const unsigned char*>(P + int64_t(k / HP) * b) + (k % HP) * RING_HALF;
  unsigned char* dst = ring + (k % RING_SLOTS) * RING_HALF;
  const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int q = 0; q < ring_issue<NT>(); ++q) {
    const int c = q * NT + wave * 64;  // first 16-byte chunk of this wave's copy
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + (c + lane) * 16),
                                     (__attribute__((address_space(3))) void*)(dst + c * 16), 16, 0, 0);
  }
}

// HP = 2: 64 KB rows in two 32 KB halves, as above.  HP = 1: 32 KB rows
// (FP64 b = 2^12, FP32 b = 2^13), one slot each, so rows t+1 .. t+4 are in
// flight while row t is gathered.  In both, the pieces of row t - 1 are
// refilled after the barrier of row t.
template <typename T, bool FWHT, int D, int NT, int HP>
__device__ __forceinline__ void decompress_ring_body(const DecompressArgs& a, int blk) {
  static_assert(NT % 64 == 0, "emit_rows sums over whole waves");
  constexpr int ER = rows_per_thread<T, D>();
  constexpr int NH = HP * D;  // pieces to copy
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = a.b;  // b * sizeof(T) == HP * 32768
  const int tiles = (a.p + NT - 1) / NT;
  const int tid = threadIdx.x;
  const int j = int(blk % tiles) * NT + tid;
  const int i0 = int(blk / tiles) * ER;
  const bool jok = j < a.p;
  const uint32_t mask = uint32_t(b - 1);
  const T* __restrict__ P = static_cast<const T*>(a.P);
  uint32_t eb[D];
#pragma unroll
  for (int t = 0; t < D; ++t) eb[t] = jok ? a.hB[int64_t(t) * a.p + j] : 0u;
  // pieces 0 .. 4 - HP first; row t's step issues pieces HP t + 5 - HP ..
  // HP t + 4 into the slots row t - 1 held
#pragma unroll
  for (int k = 0; k < (NH < RING_SLOTS - HP ? NH : RING_SLOTS - HP); ++k)
    stage_half_async<T, NT, HP>(P, smem_raw, b, k, tid);
  T v[ER][keep_count<T, D>()];
  static_for<0, D>([&](auto tc) {
    constexpr int t = decltype(tc)::value;
    // pieces 0 .. HP t + HP - 1 landed; the ones issued after them (up to
    // HP t + 4 - HP, ring_issue copies of this wave each) may be in flight
    constexpr int issued = (HP * t + RING_SLOTS - 1 - HP < NH - 1) ? HP * t + RING_SLOTS - 1 - HP : NH - 1;
    constexpr int ahead = issued - (HP * t + HP - 1) > 0 ? issued - (HP * t + HP - 1) : 0;
    static_assert(ahead * ring_issue<NT>() < 16, "vmcnt in the low four bits");
    __builtin_amdgcn_s_waitcnt(0x0070 | (ahead * ring_issue<NT>()));  // vmcnt(...) lgkmcnt(0)
    __syncthreads();  // every wave's copies landed, and row t-1 is no longer read
#pragma unroll
    for (int k = HP * t + RING_SLOTS - HP; k <= HP * t + RING_SLOTS - 1; ++k)
      if (k < NH) stage_half_async<T, NT, HP>(P, smem_raw, b, k, tid);
    // row t's pieces sit in slots HP t mod 5 (and the next): adjacent, so the
    // byte offset is lo + h sizeof(T), except where slot 4 wraps to slot 0
    const uint32_t lo = uint32_t((HP * t) % RING_SLOTS) * RING_HALF;
    const bool wraps = HP == 2 && (HP * t) % RING_SLOTS == RING_SLOTS - 1;
    const uint32_t* hA = a.hA;
    asm volatile("" : "+s"(hA));  // row t's entries are loaded here, not all hoisted to the start
    T x[ER];
#pragma unroll
    for (int e = 0; e < ER; ++e) {
      const int i = min(i0 + e, a.m - 1);
      const uint32_t ea = hash_entry(hA, int64_t(t) * a.m + i);
      const uint32_t h = FWHT ? (hbucket(ea) ^ hbucket(eb[t])) : ((hbucket(ea) + hbucket(eb[t])) & mask);
      uint32_t off = lo + h * uint32_t(sizeof(T));
      if (wraps) off = min(off, off - uint32_t(RING_LDS));  // past slot 4: back to slot 0
      const T val = *reinterpret_cast<const T*>(smem_raw + off);
      x[e] = signed_by(val, ea, eb[t]);
    }
    // formed here, after all of the row's gathers are issued: the entries
    // are not kept live until the selection, which overlaps the copies in
    // flight instead of following the last row
#pragma unroll
    for (int e = 0; e < ER; ++e) asm volatile("" : "+v"(x[e]));
#pragma unroll
    for (int e = 0; e < ER; ++e) select_push<T, D, t>(v[e], x[e]);
  });
  T med[ER];
#pragma unroll
  for (int e = 0; e < ER; ++e) med[e] = select_result<T, D>(v[e]);
  emit_rows<T, ER>(a, i0, j, jok, med);
}

template <typename T, bool FWHT, int D, int NT, int HP>
__global__ void __launch_bounds__(NT) k_decompress_ring(DecompressArgs a) {
  decompress_ring_body<T, FWHT, D, NT, HP>(a, blockIdx.x);
}

template <typename T, bool FWHT, int D>
__global__ void __launch_bounds__(LNT) k_decompress_lds(DecompressArgs a) {
  decompress_lds_body<T, FWHT, D>(a, blockIdx.x);
}

template <bool FWHT, int D>
__global__ void __launch_bounds__(NT) k_decompress_batch(const DecompressArgs* __restrict__ jobs,
                                                         const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  decompress_body<float, FWHT, D>(jobs[e.x], e.y);
}

template <bool FWHT, int D>
__global__ void __launch_bounds__(LNT) k_decompress_lds_batch(const DecompressArgs* __restrict__ jobs,
                                                              const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  decompress_lds_body<float, FWHT, D>(jobs[e.x], e.y);
}

// the ring with 512 threads; every job of a class has the same row length
template <bool FWHT, int D, int HP>
__global__ void __launch_bounds__(LNT) k_decompress_ring_batch(const DecompressArgs* __restrict__ jobs,
                                                               const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  decompress_ring_body<float, FWHT, D, LNT, HP>(jobs[e.x], e.y);
}

// any d: the value of rank d/2 by counting, O(d^2) gathers per entry
template <typename T, bool FWHT>
__global__ void __launch_bounds__(NT) k_decompress_any(DecompressArgs a) {
  const int tiles = (a.p + NT - 1) / NT;
  const int j = int(blockIdx.x % tiles) * NT + threadIdx.x;
  const int i0 = int(blockIdx.x / tiles) * ROWS;
  if (j >= a.p) return;
  const T* __restrict__ P = static_cast<const T*>(a.P);
  const uint32_t mask = uint32_t(a.b - 1);
  const int r = a.d / 2;
  const int i1 = min(a.m, i0 + ROWS);
  for (int i = i0; i < i1; ++i) {
    T med = T(0);
    for (int u = 0; u < a.d; ++u) {
      const T vu = gather<T, FWHT>(P, a.hA[int64_t(u) * a.m + i], a.hB[int64_t(u) * a.p + j],
                                   mask, int64_t(u) * a.b);
      int less = 0, leq = 0;
      for (int t = 0; t < a.d; ++t) {
        const T vt = gather<T, FWHT>(P, a.hA[int64_t(t) * a.m + i], a.hB[int64_t(t) * a.p + j],
                                     mask, int64_t(t) * a.b);
        less += vt < vu;
        leq += vt <= vu;
      }
      if (less <= r && r < leq) med = vu;
    }
    emit<T>(a, i, j, med);
  }
}

// 1024 threads leave 128 registers per thread: every ring instantiation fits
// but FP32 at d = 10, whose 12 rows x 10 values plus hash entries spill,
// FP32 at d = 13 with 32 KB rows (one register more), and the FP64 FWHT at
// d = 5 (its 12 estimates live through the recovery prefix sum)
template <typename T, bool FWHT, int D, int HP>
constexpr bool ring_wide_fits() {
  return !(sizeof(T) == 4 && (D == 10 || (D == 13 && HP == 1))) && !(sizeof(T) == 8 && FWHT && D == 5);
}

// The ring kernel with NT threads: false when the 160 KB of LDS is refused.
template <typename T, bool FWHT, int D, int NT, int HP>
bool launch_ring(const DecompressArgs& a, hipStream_t s) {
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(&k_decompress_ring<T, FWHT, D, NT, HP>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, RING_LDS) != hipSuccess)
    return false;
  constexpr int er = rows_per_thread<T, D>();
  const int64_t blocks = int64_t((a.p + NT - 1) / NT) * ((a.m + er - 1) / er);
  k_decompress_ring<T, FWHT, D, NT, HP><<<dim3(unsigned(blocks)), dim3(NT), RING_LDS, s>>>(a);
  return true;
}

// 1024 threads where the instantiation fits in 128 registers and the
// product has at least 1024 columns, else 512; a variant that does not fit
// is never instantiated
template <typename T, bool FWHT, int D, int HP>
bool launch_ring_nt(const DecompressArgs& a, hipStream_t s, bool wide) {
  if constexpr (ring_wide_fits<T, FWHT, D, HP>()) {
    if (wide) return launch_ring<T, FWHT, D, 1024, HP>(a, s);
  }
  return launch_ring<T, FWHT, D, LNT, HP>(a, s);
}

template <typename T, bool FWHT>
hipError_t launch_t(const DecompressArgs& a, hipStream_t s) {
  const size_t row_bytes = size_t(a.b) * sizeof(T);
  if (a.d <= 16 && (row_bytes == 65536 || row_bytes == 32768)) {  // the ring of 32 KB slots
    // with at least 1024 columns: 1024 threads, so each workgroup's staging
    // of the d rows serves twice the entries
    const bool wide = a.p >= 1024;
    const bool halves = row_bytes == 65536;
    switch (a.d) {
#define CM_D(N)                                                                               \
  case N:                                                                                     \
    if (halves ? launch_ring_nt<T, FWHT, N, 2>(a, s, wide) : launch_ring_nt<T, FWHT, N, 1>(a, s, wide)) \
      return hipGetLastError();                                                               \
    break;
      CM_D(1) CM_D(2) CM_D(3) CM_D(4) CM_D(5) CM_D(6) CM_D(7) CM_D(8)
      CM_D(9) CM_D(10) CM_D(11) CM_D(12) CM_D(13) CM_D(14) CM_D(15) CM_D(16)
#undef CM_D
    }
    (void)hipGetLastError();  // 160 KB refused: the two-buffer kernel below
  }
  if (a.d <= 16 && size_t(a.b) * sizeof(T) <= 65536 && size_t(a.b) * sizeof(T) >= 16) {
    const size_t lds = 2 * size_t(a.b) * sizeof(T);
    switch (a.d) {
#define CM_D(N)                                                                       \
  case N:                                                                           \
    if (lds > 65536 && hipFuncSetAttribute(reinterpret_cast<const void*>(&k_decompress_lds<T, FWHT, N>), \
                                           hipFuncAttributeMaxDynamicSharedMemorySize,  \
                                           int(lds)) != hipSuccess)                     \
      break;                                                                        \
    k_decompress_lds<T, FWHT, N><<<dim3(unsigned(int64_t((a.p + LNT - 1) / LNT) *           \
                                          ((a.m + rows_per_thread<T, N>() - 1) /         \
                                           rows_per_thread<T, N>()))),                   \
                                   dim3(LNT), lds, s>>>(a);                              \
    return hipGetLastError();
      CM_D(1) CM_D(2) CM_D(3) CM_D(4) CM_D(5) CM_D(6) CM_D(7) CM_D(8)
      CM_D(9) CM_D(10) CM_D(11) CM_D(12) CM_D(13) CM_D(14) CM_D(15) CM_D(16)
#undef CM_D
    }
    (void)hipGetLastError();  // attribute refused: use the L2-gather kernel below
  }
  const dim3 grid(unsigned(int64_t((a.p + NT - 1) / NT) * ((a.m + ROWS - 1) / ROWS)));
  const dim3 block(NT);
  switch (a.d) {
#define CM_D(N) \
  case N: k_decompress<T, FWHT, N><<<grid, block, 0, s>>>(a); break;
    CM_D(1) CM_D(2) CM_D(3) CM_D(4) CM_D(5) CM_D(6) CM_D(7) CM_D(8)
    CM_D(9) CM_D(10) CM_D(11) CM_D(12) CM_D(13) CM_D(14) CM_D(15) CM_D(16)
#undef CM_D
    default: k_decompress_any<T, FWHT><<<grid, block, 0, s>>>(a); break;
  }
  return hipGetLastError();
}

}  // namespace

namespace {
template <bool FWHT, int D, int HP>
const void* ring_batch_fn() {
  return reinterpret_cast<const void*>(&k_decompress_ring_batch<FWHT, D, HP>);
}

const void* ring_batch_kernel(int fwht, int d, int hp) {
  switch (d) {
#define CM_D(N)                                                                                \
  case N:                                                                                      \
    return fwht ? (hp == 2 ? ring_batch_fn<true, N, 2>() : ring_batch_fn<true, N, 1>())        \
                : (hp == 2 ? ring_batch_fn<false, N, 2>() : ring_batch_fn<false, N, 1>());
    CM_D(1) CM_D(2) CM_D(3) CM_D(4) CM_D(5) CM_D(6) CM_D(7) CM_D(8)
    CM_D(9) CM_D(10) CM_D(11) CM_D(12) CM_D(13) CM_D(14) CM_D(15) CM_D(16)
#undef CM_D
    default: return nullptr;
  }
}
}  // namespace

int decompress_batch_form(int fwht, int d, int b) {
  const size_t row = size_t(b) * sizeof(float);
  if (d < 1 || d > 16 || row < 16 || row > 65536) return 0;
  if (row <= 16384) return 1;
  const int hp = row == 65536 ? 2 : 1;
  // one attribute call per kernel and device: 0 not yet set, 1 set, -1 refused
  constexpr int kDevs = 16;
  static std::atomic<int> state[kDevs][2][17][3];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) dev = 0;
  std::atomic<int> local{0};  // devices past kDevs: set the attribute every time
  std::atomic<int>& st = dev < kDevs ? state[dev][fwht != 0][d][hp] : local;
  int v = st.load(std::memory_order_acquire);
  if (v == 0) {
    v = hipFuncSetAttribute(ring_batch_kernel(fwht, d, hp), hipFuncAttributeMaxDynamicSharedMemorySize,
                            RING_LDS) == hipSuccess
            ? 1
            : -1;
    (void)hipGetLastError();
    st.store(v, std::memory_order_release);
  }
  if (v > 0) return hp == 2 ? 3 : 2;
  return row <= 32768 ? 1 : 0;  // without the ring: two buffers up to 32 KB rows, else L2
}

size_t decompress_batch_lds(int form, int b) {
  if (form == 1) return 2 * size_t(b) * sizeof(float);
  return form >= 2 ? size_t(RING_LDS) : 0;
}

int decompress_batch_blocks(int d, int form, int m, int p) {
  if (m <= 0 || p <= 0) return 0;
  if (form == 0) return int(int64_t((p + NT - 1) / NT) * ((m + ROWS - 1) / ROWS));
  int er = 1;
  switch (d) {
#define CM_D(N) case N: er = rows_per_thread<float, N>(); break;
    CM_D(1) CM_D(2) CM_D(3) CM_D(4) CM_D(5) CM_D(6) CM_D(7) CM_D(8)
    CM_D(9) CM_D(10) CM_D(11) CM_D(12) CM_D(13) CM_D(14) CM_D(15) CM_D(16)
#undef CM_D
    default: return -1;
  }
  return int(int64_t((p + LNT - 1) / LNT) * ((m + er - 1) / er));
}

hipError_t launch_decompress_batch(int fwht, int d, int form, size_t lds,
                                   const DecompressArgs* jobs, const int2* blocks, int nblocks,
                                   hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  if (d < 1 || d > 16 || form < 0 || form > 3 || (form == 1 && lds > 65536) ||
      (form >= 2 && lds != size_t(RING_LDS)))
    return hipErrorInvalidValue;
  const dim3 grid = dim3(unsigned(nblocks));
  switch (d) {
#define CM_D(N)                                                                                   \
  case N:                                                                                         \
    if (form == 3) {                                                                              \
      if (fwht) k_decompress_ring_batch<true, N, 2><<<grid, dim3(LNT), lds, s>>>(jobs, blocks);   \
      else k_decompress_ring_batch<false, N, 2><<<grid, dim3(LNT), lds, s>>>(jobs, blocks);       \
    } else if (form == 2) {                                                                       \
      if (fwht) k_decompress_ring_batch<true, N, 1><<<grid, dim3(LNT), lds, s>>>(jobs, blocks);   \
      else k_decompress_ring_batch<false, N, 1><<<grid, dim3(LNT), lds, s>>>(jobs, blocks);       \
    } else if (form == 1) {                                                                       \
      if (fwht) k_decompress_lds_batch<true, N><<<grid, dim3(LNT), lds, s>>>(jobs, blocks);       \
      else k_decompress_lds_batch<false, N><<<grid, dim3(LNT), lds, s>>>(jobs, blocks);           \
    } else {                                                                                      \
      if (fwht) k_decompress_batch<true, N><<<grid, dim3(NT), 0, s>>>(jobs, blocks);              \
      else k_decompress_batch<false, N><<<grid, dim3(NT), 0, s>>>(jobs, blocks);                  \
    }                                                                                             \
    break;
    CM_D(1) CM_D(2) CM_D(3) CM_D(4) CM_D(5) CM_D(6) CM_D(7) CM_D(8)
    CM_D(9) CM_D(10) CM_D(11) CM_D(12) CM_D(13) CM_D(14) CM_D(15) CM_D(16)
#undef CM_D
  }
  return hipGetLastError();
}

hipError_t launch_decompress(const DecompressArgs& a, hipStream_t s) {
  if (a.m == 0 || a.p == 0) return hipSuccess;
  if (a.f64) return a.fwht ? launch_t<double, true>(a, s) : launch_t<double, false>(a, s);
  return a.fwht ? launch_t<float, true>(a, s) : launch_t<float, false>(a, s);
}

}  // namespace cm
