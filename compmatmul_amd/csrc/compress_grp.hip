// Slice-group FP32 FWHT compress: one scatter per row for G slices.
//
// The single-wave kernel (compress_reg.hip) gives every slice of the
// spectrum its own modulated count-sketch, so each row is scattered b1
// times.  Here a workgroup of G = 2^LOG_G waves owns G consecutive slices
// s = group*G + w of one repetition:
//  * the row is staged once into LDS (2048 keys at a time) by all waves;
//  * it is scattered once into a G x 2048-bucket sketch S[hw][l] (h = hash,
//    l = h mod 2048, hw = (h >> 11) mod G), modulated by the group bits
//    (-1)^popc((h >> (11 + LOG_G)) & group).  Wave w scatters only the keys
//    with hw = w, into its own partition, following a conflict-free schedule
//    (rounds by occurrence within the bucket, as in compress_reg.hip);
//  * a G-point Walsh-Hadamard transform across the partitions,
//      x_s[l] = sum_hw (-1)^popc(hw & w) S[hw][l],
//    turns partition w into the modulated sketch of slice s (the sign
//    (-1)^popc((h >> 11) & s) splits into the hw and group parts);
//  * each wave then runs the 2048-point register WHT of compress_reg.hip on
//    its partition and keeps its slice's accumulator in 32 VGPRs.
// Per (t, k) this scatters each key once per G slices instead of G times,
// and reads each row once per group; the partial spectra keep the
// [chunk][t][slice][2048] layout, so reduction and combine are unchanged.
#include "device_common.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

constexpr int LB2 = 11;          // slice length 2048
constexpr int B2 = 1 << LB2;
constexpr int R = B2 / 64;       // values per lane (32)
constexpr int TROW = 33;         // padded transpose row
// Floats per partition row and per partition: G = 4 pads rows of 32 to 36
// floats (16-byte rows, read back with ds_read_b128), G <= 2 to 33.
template <int LOG_G>
__host__ __device__ constexpr int prow() {
  return LOG_G == 2 ? 36 : TROW;
}
template <int LOG_G>
__host__ __device__ constexpr int plen() {
  return 64 * prow<LOG_G>();
}
constexpr int SEG = 2048;        // keys staged at a time
constexpr int RCAP = 8;
constexpr int MAXSEG = 32;       // keys <= MAXSEG * SEG = 65536
constexpr uint32_t F_SIGN = 1u << 31;
constexpr int GB = 8;            // schedule groups per batch

__device__ __forceinline__ float flip(float v, uint32_t sgn) {
  return __int_as_float(__float_as_int(v) ^ sgn);
}

// Partition layout.  G = 4 stores position l of a partition at
// (l >> 5) * 36 + (l & 31): 64 rows of 32 positions padded to 36 floats, so
// the in-place write of the cross pass is conflict-free per half-wave and
// the transposed read is eight 16-byte reads per lane, each pass of 8 lanes
// on 32 distinct banks; slot 32 of row i is the dummy slot of lane i.  G <= 2
// keeps positions in order, dummies at 2048 + lane.
template <int LOG_G>
__host__ __device__ __forceinline__ uint32_t slot_of(uint32_t l) {
  return LOG_G == 2 ? (l >> 5) * prow<2>() + (l & 31) : l;
}
template <int LOG_G>
__host__ __device__ __forceinline__ uint32_t dummy_slot(int lane) {
  return LOG_G == 2 ? uint32_t(lane) * prow<2>() + 32 : uint32_t(B2 + lane);
}

__host__ __device__ __forceinline__ int64_t list_cap(int keys) {
  const int64_t segs = (int64_t(keys) + SEG - 1) / SEG;
  return int64_t(keys) + segs * 64 * RCAP;
}

struct Layout {
  uint32_t* LA;
  uint32_t* LB;
  int* ngroups;
  int64_t cap_a, cap_b;
  int lists;  // d * b1: one per (t, slice) and operand
};

__host__ __device__ __forceinline__ Layout layout(const RegJob& j) {
  Layout o;
  o.lists = j.d << j.lbs;
  o.cap_a = list_cap(j.m);
  o.cap_b = list_cap(j.p);
  o.LA = static_cast<uint32_t*>(j.lists);
  o.LB = o.LA + o.lists * o.cap_a;
  o.ngroups = reinterpret_cast<int*>(o.LB + o.lists * o.cap_b);
  return o;
}

// Schedule of list g (g < lists: A's row for slice ts = g, else B's): the
// keys of partition w = slice mod G, bucket l = h mod 2048, sign modulated
// by the group bits.  Entry: byte offset of the destination in the
// workgroup's partitions, (w plen + slot_of(l)) << 2 (or w plen + the dummy
// slot of the lane; below 2^16) | key offset in the 2048-key segment << 18
// | sign << 31.
// LW waves build one list: wave bw takes keys [bw SEG/LW, (bw+1) SEG/LW)
// of each segment.  A key's round is its occurrence index among the keys of
// its bucket in key order: the count from earlier segments, plus the counts
// of the earlier waves, plus its rank in its own wave (one wave's LDS
// atomics complete in program order).  Within a round, keys are placed in
// key order: each wave's run starts after the earlier waves' keys of that
// round.  The list is the same as one wave walking the keys in order.
constexpr int LW = 4;

template <int LOG_G>
__device__ __forceinline__ void grp_lists_body(const RegJob& j, int g) {
  constexpr int G = 1 << LOG_G;
  constexpr int KW = SEG / LW;   // keys per wave and segment
  constexpr int Q = KW / 64;     // keys per lane and segment
  __shared__ uint32_t cnt[B2];     // occurrences in earlier segments
  __shared__ uint32_t cw[LW][B2];  // occurrences in this segment, per wave
  __shared__ int nrw[LW][RCAP];    // this segment's keys per wave and round
  const Layout lay = layout(j);
  const bool side_b = g >= lay.lists;
  const int ts = side_b ? g - lay.lists : g;
  const int t = ts >> j.lbs;
  const uint32_t slice = uint32_t(ts & ((1 << j.lbs) - 1));
  const uint32_t w = slice & (G - 1), group = slice >> LOG_G;
  const uint32_t pbase = w * plen<LOG_G>();
  static_assert((plen<LOG_G>() << (LOG_G + 2)) <= (1 << 16), "destinations fit 16 bits");
  const int keys = side_b ? j.p : j.m;
  const uint32_t* __restrict__ hh = (side_b ? j.hB : j.hA) + int64_t(t) * keys;
  uint32_t* __restrict__ L = side_b ? lay.LB + ts * lay.cap_b : lay.LA + ts * lay.cap_a;
  const int tid = threadIdx.x;
  const int bw = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  for (int e = tid; e < B2; e += 64 * LW) cnt[e] = 0;
  for (int e = tid; e < LW * B2; e += 64 * LW) (&cw[0][0])[e] = 0;
  int out = 0;
  for (int s0 = 0; s0 < keys; s0 += SEG) {
    const int sl = min(SEG, keys - s0);
    // this wave's hash entries of the segment in flight at once
    uint32_t ek[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int kk = bw * KW + q * 64 + lane;
      ek[q] = kk < sl ? hh[s0 + kk] : 0u;
    }
    if (lane < RCAP) nrw[bw][lane] = 0;
    __syncthreads();
    int loc[Q];  // rank within this wave, -1 outside this partition
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const uint32_t h = hbucket(ek[q]);
      loc[q] = -1;
      if (bw * KW + q * 64 + lane < sl && ((h >> LB2) & (G - 1)) == w)
        loc[q] = int(atomicAdd(&cw[bw][h & (B2 - 1)], 1u));
    }
    __syncthreads();
    int rk[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      rk[q] = -1;
      if (loc[q] >= 0) {
        const uint32_t b = hbucket(ek[q]) & (B2 - 1);
        int r = int(cnt[b]) + loc[q];
        for (int w2 = 0; w2 < bw; ++w2) r += int(cw[w2][b]);
        rk[q] = min(r, RCAP - 1);
        atomicAdd(&nrw[bw][rk[q]], 1);
      }
    }
    __syncthreads();
    static_assert(RCAP == 8, "round_slot places rounds 0..7");
    int base[RCAP], tot[RCAP];
    int64_t run[RCAP];  // next slot of each round for this wave
    int acc = out;
#pragma unroll
    for (int r = 0; r < RCAP; ++r) {
      int before = 0, all = 0;
#pragma unroll
      for (int w2 = 0; w2 < LW; ++w2) {
        const int n = nrw[w2][r];
        before += w2 < bw ? n : 0;
        all += n;
      }
      base[r] = acc;
      tot[r] = all;
      run[r] = int64_t(acc) * 64 + before;
      acc += (all + 63) / 64;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = rk[q];
      const int64_t pos = round_slot(r, lane, run);
      if (r >= 0) {
        const uint32_t h = hbucket(ek[q]);
        const uint32_t neg = hneg(ek[q]) ^ (__popc((h >> (LB2 + LOG_G)) & group) & 1u);
        L[pos] = ((pbase + slot_of<LOG_G>(h & (B2 - 1))) << 2) | (uint32_t(bw * KW + q * 64 + lane) << 18) |
                 (neg ? F_SIGN : 0u);
      }
    }
#pragma unroll
    for (int rr = 0; rr < RCAP; ++rr) {  // pad rounds to whole groups (dummy slots)
      const int full = (tot[rr] + 63) / 64 * 64;
      for (int slot = tot[rr] + tid; slot < full; slot += 64 * LW)
        L[int64_t(base[rr]) * 64 + slot] = (pbase + dummy_slot<LOG_G>(slot & 63)) << 2;
    }
    if (tid < RCAP) {
      int all = 0;
#pragma unroll
      for (int w2 = 0; w2 < LW; ++w2) all += nrw[w2][tid];
      const int64_t row = int64_t(g) * MAXSEG + s0 / SEG;
      lay.ngroups[row * RCAP + tid] = (all + 63) / 64;
    }
    out = acc;
    __syncthreads();  // every rank and count of the segment is read
    for (int e = tid; e < B2; e += 64 * LW) {  // fold the segment into cnt
      uint32_t c = cnt[e];
#pragma unroll
      for (int w2 = 0; w2 < LW; ++w2) {
        c += cw[w2][e];
        cw[w2][e] = 0;
      }
      cnt[e] = c;
    }
  }
}

template <int LOG_G>
__global__ void __launch_bounds__(64 * LW) k_grp_lists(RegJob j) { grp_lists_body<LOG_G>(j, blockIdx.x); }

template <int LOG_G>
__global__ void __launch_bounds__(64 * LW) k_grp_lists_batch(const RegJob* __restrict__ jobs,
                                                          const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  grp_lists_body<LOG_G>(jobs[e.x], e.y);
}

// x[i] = x[i] of lane L ^ 1 times s, plus x[i], for i < 8: one
// v_fmac_f32 with a DPP quad_perm source each (the compiler does not fold a
// DPP move into an fma).  The s_nop gives the two wait states a DPP read of
// a VGPR needs after a VALU write; each x[i] is read once, before its write.
__device__ __forceinline__ void dpp_butterfly8(float* x, float s) {
  asm volatile(
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %0, %8 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %1, %1, %8 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %2, %2, %8 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %3, %3, %8 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %4, %4, %8 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %5, %5, %8 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %6, %6, %8 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %7, %7, %8 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf"
      : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]), "+v"(x[4]), "+v"(x[5]), "+v"(x[6]), "+v"(x[7])
      : "v"(s));
}

// As dpp_butterfly8 with partner lane L ^ 2 (quad_perm [2, 3, 0, 1]).
__device__ __forceinline__ void dpp_butterfly8_x2(float* x, float s) {
  asm volatile(
      "s_nop 1\n\t"
      "v_fmac_f32_dpp %0, %0, %8 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %1, %1, %8 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %2, %2, %8 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %3, %3, %8 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %4, %4, %8 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %5, %5, %8 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %6, %6, %8 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f32_dpp %7, %7, %8 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf"
      : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]), "+v"(x[4]), "+v"(x[5]), "+v"(x[6]), "+v"(x[7])
      : "v"(s));
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

// KIND 0 plain writes, 1 read-add-writes, 2 LDS atomics (see compress_reg.hip)
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
      v[q] = flip(*reinterpret_cast<const float*>(Rbb + ((ent[q] >> 16) & 0x1ffcu)), ent[q] & F_SIGN);
#pragma unroll
    for (int q = 0; q < GB; ++q) {
      if (q < nb) {
        float* d = reinterpret_cast<float*>(Sb + (ent[q] & 0xfffcu));
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

// Schedule groups kept in registers for a whole unit: the schedule of a
// (t, slice) list does not depend on k.  Slots past the list hold dummy
// entries (plain writes of Rb[0] into the lane's dummy slot of partition w).
constexpr int CA = 12;

// An empty operand (keys == 0) has no schedule: its counts are never
// written, so they are not read.
template <int LOG_G>
__device__ __forceinline__ void load_cache(const uint32_t* __restrict__ lp,
                                           const int* __restrict__ nr, int keys, uint32_t (&c)[CA],
                                           int w, int lane) {
  int ng = 0;
  if (keys > 0)
    for (int r = 0; r < RCAP; ++r) ng += nr[r];
#pragma unroll
  for (int g = 0; g < CA; ++g)
    c[g] = g < ng ? lp[g * 64 + lane] : (w * plen<LOG_G>() + dummy_slot<LOG_G>(lane)) << 2;
}

// Section ends of a single-segment list in groups: plain writes below n0,
// read-add-writes below e1, atomics below ng.  They depend on the list only,
// so they are read once per unit, into scalar registers.
struct GroupCounts {
  int n0, e1, ng;
};

__device__ __forceinline__ GroupCounts group_counts(const int* __restrict__ nr, int keys) {
  GroupCounts gc{0, 0, 0};
  if (keys > 0) {
    gc.n0 = __builtin_amdgcn_readfirstlane(nr[0]);
    int e1 = gc.n0;
    for (int r = 1; r < RCAP - 1; ++r) e1 += nr[r];
    gc.e1 = __builtin_amdgcn_readfirstlane(e1);
    gc.ng = __builtin_amdgcn_readfirstlane(e1 + nr[RCAP - 1]);
  }
  return gc;
}

// The scatter of one single-segment list from the cached groups, then any
// groups past the cache from memory.  Group g is a plain write below nr[0],
// a read-add-write below nr[0] + (rounds 1 .. RCAP-2), else an LDS atomic;
// one wave's LDS operations complete in order, as the rounds require.
__device__ __forceinline__ void scatter_cached(uint32_t (&c)[CA],
                                               const uint32_t* __restrict__ lp,
                                               const GroupCounts& gc, float* S,
                                               const float* Rb, int lane) {
  // opaque here, so each group's tests stay scalar compares in the loop
  // instead of being hoisted out as 24 live condition masks
  int n0 = gc.n0, e1 = gc.e1, ng = gc.ng;
  asm volatile("" : "+s"(n0), "+s"(e1), "+s"(ng));
  const char* Rbb = reinterpret_cast<const char*>(Rb);
  char* Sb = reinterpret_cast<char*>(S);
  constexpr int BATCH = 6;  // LDS reads in flight per batch (4: 1% slower)
#pragma unroll
  for (int g0 = 0; g0 < CA; g0 += BATCH) {
    // a batch wholly past the list is skipped (one scalar branch: short
    // lists of small products); within a batch the reads stay together
    if (g0 >= ng) break;
    // the entries are made opaque in place, so their fields are extracted
    // here rather than hoisted out of the k loop as three live registers
    uint32_t e[BATCH];
    float v[BATCH];
#pragma unroll
    for (int q = 0; q < BATCH; ++q) {
      asm volatile("" : "+v"(c[g0 + q]));
      e[q] = c[g0 + q];
      v[q] = flip(*reinterpret_cast<const float*>(Rbb + ((e[q] >> 16) & 0x1ffcu)), e[q] & F_SIGN);
    }
#pragma unroll
    for (int q = 0; q < BATCH; ++q) {
      const int g = g0 + q;
      float* d = reinterpret_cast<float*>(Sb + (e[q] & 0xfffcu));
      if (g < n0 || g >= ng)
        *d = v[q];
      else if (g < e1)
        *d += v[q];
      else
        __hip_atomic_fetch_add(d, v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  if (ng > CA) {  // a deep schedule: the rest from memory, section by section
    if (n0 > CA) scatter_section<0>(lp + CA * 64, n0 - CA, S, Rb, lane);
    const int s1 = max(n0, CA);
    if (e1 > s1) scatter_section<1>(lp + int64_t(s1) * 64, e1 - s1, S, Rb, lane);
    const int s2 = max(e1, CA);
    if (ng > s2) scatter_section<2>(lp + int64_t(s2) * 64, ng - s2, S, Rb, lane);
  }
}

// The last WHT stage pairs lane L with lane L ^ 32: v_permlane32_swap brings
// each pair (p, q) into registers 2jj, 2jj + 1 of one lane.  Its butterfly
// (p + q, p - q) is not applied: the product of the A and B pairs is
// accumulated in bilinear form instead (accumulate_pairs), four fmas per
// pair where transforming both operands and multiplying takes six
// operations.
__device__ __forceinline__ void swap_halves(float (&x)[R]) {
#pragma unroll
  for (int jj = 0; jj < R / 2; ++jj) {
    const auto pq = __builtin_amdgcn_permlane32_swap(__float_as_uint(x[2 * jj]),
                                                     __float_as_uint(x[2 * jj + 1]), false, false);
    x[2 * jj] = __uint_as_float(pq[0]);
    x[2 * jj + 1] = __uint_as_float(pq[1]);
  }
}

// For pairs (pa, qa) and (pb, qb) before their last butterfly,
// (pa + qa)(pb + qb) = Sum + Dif and (pa - qa)(pb - qb) = Sum - Dif with
// Sum = pa pb + qa qb and Dif = pa qb + qa pb: acc[2jj] accumulates Sum and
// acc[2jj + 1] Dif, and finish_pairs turns them into the two products.
__device__ __forceinline__ void accumulate_pairs(const float (&a)[R], const float (&b)[R],
                                                 float (&acc)[R]) {
#pragma unroll
  for (int r = 0; r < R; r += 2) {
    acc[r] = fmaf(a[r + 1], b[r + 1], fmaf(a[r], b[r], acc[r]));
    acc[r + 1] = fmaf(a[r + 1], b[r], fmaf(a[r], b[r + 1], acc[r + 1]));
  }
}

__device__ __forceinline__ void finish_pairs(float (&acc)[R]) {
#pragma unroll
  for (int r = 0; r < R; r += 2) {
    const float sum = acc[r], dif = acc[r + 1];
    acc[r] = sum + dif;
    acc[r + 1] = sum - dif;
  }
}

// Scatter one row for the G slices of the group, cross-partition WHT, then
// the register WHT of this wave's slice.  On return lane L holds slice
// group*G + w with every butterfly applied but the last: the butterfly of
// registers (r, r + 1), r even, gives frequencies f_of(L, r) and
// f_of(L, r + 1) (f_of4 for G = 4).
// With `staged`, the row (one segment) is already in Rb; `after_scatter`
// runs once every wave has finished reading it.
template <int LOG_G, bool CACHED, typename AfterScatter>
__device__ __forceinline__ void sketch_transform_grp(const float* __restrict__ row, int keys,
                                                     const uint32_t* __restrict__ list,
                                                     const int* __restrict__ segcnt, float* S,
                                                     float* Rb, bool staged, int w, int lane,
                                                     float (&x)[R], uint32_t (&cache)[CA],
                                                     const GroupCounts& gc,
                                                     AfterScatter after_scatter, bool drain) {
  constexpr int G = 1 << LOG_G;
  float* Sw = S + w * plen<LOG_G>();
  const int tid = w * 64 + lane;
  if constexpr (LOG_G == 2) {
    // the 32 data columns of the 64 padded rows (the pad columns hold only
    // dummy slots, never read): lane L clears columns 4 (L & 7) .. + 3 of
    // rows 8 q + (L >> 3)
    float4* z = reinterpret_cast<float4*>(Sw + (lane >> 3) * prow<2>() + (lane & 7) * 4);
#pragma unroll
    for (int q = 0; q < 8; ++q) z[q * 8 * prow<2>() / 4] = make_float4(0.f, 0.f, 0.f, 0.f);
  } else {
    float4* S4 = reinterpret_cast<float4*>(Sw);
#pragma unroll
    for (int q = 0; q < B2 / 4 / 64; ++q) S4[q * 64 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  const int nseg = (keys + SEG - 1) / SEG;
  const uint32_t* __restrict__ lp = list;
  if constexpr (CACHED) scatter_cached(cache, list, gc, S, Rb, lane);  // staged, one segment
  for (int seg = 0; seg < (CACHED ? 0 : nseg); ++seg) {
    if (!staged) {
      const int s0 = seg * SEG;
      const int sl = min(SEG, keys - s0);
      float v[SEG / (64 * G)];
#pragma unroll
      for (int q = 0; q < SEG / (64 * G); ++q) v[q] = row[s0 + min(q * 64 * G + tid, sl - 1)];
      __syncthreads();  // every wave is done with the previous segment
#pragma unroll
      for (int q = 0; q < SEG / (64 * G); ++q) Rb[q * 64 * G + tid] = v[q];
      __syncthreads();
    }
    const int* __restrict__ nr = segcnt + seg * RCAP;
    int nw = 0;
    for (int r = 1; r < RCAP - 1; ++r) nw += nr[r];
    const int n0 = __builtin_amdgcn_readfirstlane(nr[0]);  // uniform: scalar loop control
    const int n1 = __builtin_amdgcn_readfirstlane(nw);
    const int n2 = __builtin_amdgcn_readfirstlane(nr[RCAP - 1]);
    scatter_section<0>(lp, n0, S, Rb, lane);
    lp += int64_t(n0) * 64;
    scatter_section<1>(lp, n1, S, Rb, lane);
    lp += int64_t(n1) * 64;
    scatter_section<2>(lp, n2, S, Rb, lane);
    lp += int64_t(n2) * 64;
  }
  __syncthreads();  // all partitions complete
  after_scatter();   // the staged row is free again
  if constexpr (LOG_G == 2) {
    // One register pass does the cross-partition WHT together with
    // position bits 0..2: lane L, register (c << 2) | hw holds position p
    // of partition hw, with p bits 0..2 = c, bits 3, 4, 5 = L bits 0, 1, 2,
    // bit 6 = w bit 0, bits 7, 8, 9 = L bits 3, 4, 5, bit 10 = w bit 1 (two
    // 16-byte reads per partition: the eight positions share a padded row,
    // and each pass of 8 lanes covers 32 distinct banks).  The results go
    // back into this wave's region (the rows whose bits 1 and 5 are w's,
    // which no other wave reads), permuted for the transposed read: register
    // (c << 2) | j to row (L >> 5) + 2 (w & 1) + 4 c + 32 (w >> 1), column
    // L & 31 of slice j.
    const int p0 = (lane & 3) * 8 + ((lane >> 2) & 1) * 32 + (w & 1) * 64 + ((lane >> 3) & 7) * 128 +
                   (w >> 1) * 1024;
    {
      const int a = int(slot_of<2>(uint32_t(p0)));
#pragma unroll
      for (int hw = 0; hw < 4; ++hw) {
        const float4* src = reinterpret_cast<const float4*>(S + hw * plen<2>() + a);
        const float4 u = src[0], v = src[1];
        x[(0 << 2) | hw] = u.x;
        x[(1 << 2) | hw] = u.y;
        x[(2 << 2) | hw] = u.z;
        x[(3 << 2) | hw] = u.w;
        x[(4 << 2) | hw] = v.x;
        x[(5 << 2) | hw] = v.y;
        x[(6 << 2) | hw] = v.z;
        x[(7 << 2) | hw] = v.w;
      }
    }
    wht_reg(x);
    {
      const int wb = ((lane >> 5) + 2 * (w & 1) + 32 * (w >> 1)) * prow<2>() + (lane & 31);
#pragma unroll
      for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) S[jj * plen<2>() + wb + 4 * c * prow<2>()] = x[(c << 2) | jj];
    }
    // with `drain`, this wave's row copies have landed too, so the barrier
    // also publishes the next rows
    if (drain) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();  // every wave's region is written
    // wave w now reads its slice: row `lane` holds positions l3, l4, l5,
    // l7, l8 of (l9, l6, f0, f1, f2, l10) = lane bits 0..5
    {
      const float4* rbase = reinterpret_cast<const float4*>(Sw + lane * prow<2>());
#pragma unroll
      for (int c = 0; c < R / 4; ++c) {
        const float4 v = rbase[c];
        x[4 * c] = v.x;
        x[4 * c + 1] = v.y;
        x[4 * c + 2] = v.z;
        x[4 * c + 3] = v.w;
      }
    }
    wht_reg(x);  // position bits 3, 4, 5, 7, 8
    // position bit 9 (lane bit 0): x[L ^ 1] s + x[L] with s = -1 on odd
    // lanes, which therefore hold the negated value; likewise position bit
    // 6 (lane bit 1) with x[L ^ 2].  A and B carry the same signs, so their
    // product is unchanged
    const float sgn = (lane & 1) ? -1.f : 1.f;
#pragma unroll
    for (int r = 0; r < R; r += 8) dpp_butterfly8(x + r, sgn);
    const float sgn2 = (lane & 2) ? -1.f : 1.f;
#pragma unroll
    for (int r = 0; r < R; r += 8) dpp_butterfly8_x2(x + r, sgn2);
    // position bit 10 (lane bit 5): the swap brings each pair into one lane;
    // its butterfly is left to the caller (lanes L and L ^ 32 have the same
    // parity, so the pair carries one sign)
    swap_halves(x);
    return;
  }
  if constexpr (G > 1) {
    // wave w transforms positions l in [w 2048/G, (w+1) 2048/G) across the
    // G partitions (natural-order WHT over hw)
    constexpr int PER = R / G;
    float c[G][PER];
    const int l0 = w * (B2 / G);
#pragma unroll
    for (int hw = 0; hw < G; ++hw)
#pragma unroll
      for (int q = 0; q < PER; ++q) c[hw][q] = S[hw * plen<LOG_G>() + l0 + q * 64 + lane];
#pragma unroll
    for (int st = 1; st < G; st <<= 1)
#pragma unroll
      for (int hw = 0; hw < G; ++hw) {
        if (hw & st) continue;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const float a = c[hw][q], b = c[hw | st][q];
          c[hw][q] = a + b;
          c[hw | st][q] = a - b;
        }
      }
#pragma unroll
    for (int hw = 0; hw < G; ++hw)
#pragma unroll
      for (int q = 0; q < PER; ++q) S[hw * plen<LOG_G>() + l0 + q * 64 + lane] = c[hw][q];
    if (drain) __builtin_amdgcn_s_waitcnt(0);  // as above
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = Sw[r * 64 + lane];
  // the register WHT and transposes touch only this wave's partition: LDS
  // operations of one wave complete in order, so only the compiler must be
  // kept from reordering them
  asm volatile("" ::: "memory");
  wht_reg(x);  // index bits 6..10
  {
    float* wbase = Sw + (lane >> 5) * TROW + (lane & 31);
#pragma unroll
    for (int r = 0; r < R; ++r) wbase[r * 2 * TROW] = x[r];
  }
  asm volatile("" ::: "memory");
  {
    const float* rbase = Sw + (((lane & 31) << 1) | (lane >> 5)) * TROW;
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = rbase[r];
  }
  asm volatile("" ::: "memory");
  wht_reg(x);  // index bits 0..4
  swap_halves(x);  // index bit 5 across lane halves; butterfly left to the caller
}

__device__ __forceinline__ int f_of(int L, int r) {
  return ((L & 31) << 6) | ((r & 1) << 5) | ((r & ~1) + (L >> 5));
}

// G = 4: frequency of lane L, register r after the merged cross pass
// (register bits 0..4 -> f10, f4, f5, f7, f8; lane bits 0..5 -> f9, f6,
// f0, f1, f2, f3)
__device__ __forceinline__ int f_of4(int L, int r) {
  return ((r & 1) << 10) | (((r >> 1) & 1) << 4) | (((r >> 2) & 1) << 5) | (((r >> 3) & 1) << 7) |
         (((r >> 4) & 1) << 8) | ((L & 1) << 9) | (((L >> 1) & 1) << 6) | ((L >> 2) & 7) |
         (((L >> 5) & 1) << 3);
}

// Start copying row[0, len) (len <= SEG) into LDS dst[0, SEG) with
// global_load_lds, 4 bytes per lane and no registers: wave w (uniform) of
// the workgroup copies elements q 64G + 64 w + lane; lanes past the row copy
// its last element into slots the schedule never reads.  A full row needs no
// clamp: each copy is the scalar base row + q 64G + 64 w and the lane's
// offset, with no vector arithmetic.
template <int G>
__device__ __forceinline__ void copy_row_async(const float* __restrict__ row, int len, float* dst, int w,
                                               int lane) {
  if (len <= 0) return;
  if (len == SEG) {
#pragma unroll
    for (int q = 0; q < SEG / (64 * G); ++q) {
      const int e0 = q * 64 * G + w * 64;
      const float* rq = row + e0;  // uniform; the unsigned lane offset allows scalar-base addressing
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(rq + uint32_t(lane)),
                                       (__attribute__((address_space(3))) void*)(dst + e0), 4, 0, 0);
    }
    return;
  }
  // opaque per call, so the clamped offsets are recomputed here (two VALU
  // each) instead of being hoisted out of the k loop and kept live
  int lane_id = w * 64 + lane;
  asm volatile("" : "+v"(lane_id));
#pragma unroll
  for (int q = 0; q < SEG / (64 * G); ++q) {
    const int e = q * 64 * G + lane_id;
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(row + min(e, len - 1)),
                                     (__attribute__((address_space(3))) void*)(dst + q * 64 * G + w * 64), 4,
                                     0, 0);
  }
}

// unit u = (chunk, t, group); wave w owns slice group * G + w.  PAIRED: both
// rows fit one segment (m, p <= SEG); they are copied into LDS ahead of use
// (row k + 1 of A while A's row k is transformed, likewise for B), so the
// load latency is hidden behind the transforms.
template <int LOG_G, bool PAIRED>
__device__ __forceinline__ void grp_unit(const RegJob& j, int u, float* S, float* Rb) {
  constexpr int G = 1 << LOG_G;
  const Layout lay = layout(j);
  const int per_chunk = lay.lists >> LOG_G;
  const int chunk = u / per_chunk;
  const int tg = u - chunk * per_chunk;  // t * (b1 / G) + group
  // the wave index is uniform: in a scalar register, the schedule counts
  // derived from it are scalar and each group's kind is a scalar branch
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  static_assert(LOG_G >= 1, "the cross pass barrier publishes the rows of k + 1");
  const int ts = tg * G + w;              // t * b1 + slice
  const int k0 = chunk * j.kper;
  const int k1 = min(j.n, k0 + j.kper);
  const uint32_t* __restrict__ la = lay.LA + ts * lay.cap_a;
  const uint32_t* __restrict__ lb = lay.LB + ts * lay.cap_b;
  const int* __restrict__ sca = lay.ngroups + int64_t(ts) * MAXSEG * RCAP;
  const int* __restrict__ scb = lay.ngroups + int64_t(lay.lists + ts) * MAXSEG * RCAP;
  float acc[R];
#pragma unroll
  for (int r = 0; r < R; ++r) acc[r] = 0.f;
  constexpr bool paired = PAIRED;
  // paired lists are single-segment: keep their first groups in registers
  uint32_t ca[CA], cb[CA];
  GroupCounts ga{0, 0, 0}, gb{0, 0, 0};
  if constexpr (paired) {
    load_cache<LOG_G>(la, sca, j.m, ca, w, lane);
    load_cache<LOG_G>(lb, scb, j.p, cb, w, lane);
    ga = group_counts(sca, j.m);
    gb = group_counts(scb, j.p);
  }
  if (paired && k0 < k1) {
    copy_row_async<G>(j.AT + int64_t(k0) * j.ld_at, j.m, Rb, w, lane);
    copy_row_async<G>(j.B + int64_t(k0) * j.ld_b, j.p, Rb + SEG, w, lane);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of rows k0 have landed
    __syncthreads();                // and every wave's
  }
  for (int k = k0; k < k1; ++k) {
    float xa[R], xb[R];
    const float* arow = j.AT + int64_t(k) * j.ld_at;
    const float* brow = j.B + int64_t(k) * j.ld_b;
    // rows k + 1 are copied during k: A's after A's scatter, B's after B's;
    // B's cross pass drains both before its barrier, which thereby also
    // publishes them for the scatters of k + 1
    const bool more = paired && k + 1 < k1;
    sketch_transform_grp<LOG_G, paired>(arow, j.m, la, sca, S, Rb, paired, w, lane, xa, ca, ga, [&] {
      if (more) copy_row_async<G>(arow + j.ld_at, j.m, Rb, w, lane);
    }, false);
    sketch_transform_grp<LOG_G, paired>(brow, j.p, lb, scb, S, paired ? Rb + SEG : Rb, paired, w, lane, xb,
                                        cb, gb, [&] {
                                          if (more) copy_row_async<G>(brow + j.ld_b, j.p, Rb + SEG, w, lane);
                                        }, more);
    accumulate_pairs(xa, xb, acc);
  }
  finish_pairs(acc);
  float* Sw = S + w * plen<LOG_G>();
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) Sw[LOG_G == 2 ? f_of4(lane, r) : f_of(lane, r)] = acc[r];
  __syncthreads();
  float* out = j.partial + (int64_t(chunk) * lay.lists + ts) * B2;
#pragma unroll
  for (int r = 0; r < R; ++r) out[r * 64 + lane] = Sw[r * 64 + lane];
}

// Three waves per SIMD for G = 4 (160 VGPRs, no spill; LDS also allows
// three workgroups per CU); G = 2 needs more registers per wave and targets
// two.
template <int LOG_G, bool PAIRED>
__global__ void __launch_bounds__(64 << LOG_G) __attribute__((amdgpu_waves_per_eu(LOG_G == 2 ? 3 : 2, LOG_G == 2 ? 3 : 2)))
k_compress_fwht_grp(RegJob j, int units) {
  __shared__ __align__(16) float S[plen<LOG_G>() << LOG_G];
  __shared__ __align__(16) float Rb[2 * SEG];
  const int G = gridDim.x;
  const int g = blockIdx.x;
  const int u = (g & 7) * (G >> 3) + (g >> 3);  // XCD-aware numbering
  if (u >= units) return;
  grp_unit<LOG_G, PAIRED>(j, u, S, Rb);
}

// Persistent form for many products: the XCD work queues of the single-wave
// batch kernels; thread 0 claims each item and the workgroup reads it from
// LDS.
template <int LOG_G>
__global__ void __launch_bounds__(64 << LOG_G) __attribute__((amdgpu_waves_per_eu(LOG_G == 2 ? 3 : 2, LOG_G == 2 ? 3 : 2)))
k_compress_fwht_grp_batch(const RegJob* __restrict__ jobs, const int2* __restrict__ items,
                  const int* __restrict__ qstart, unsigned* __restrict__ counters) {
  __shared__ __align__(16) float S[plen<LOG_G>() << LOG_G];
  __shared__ __align__(16) float Rb[2 * SEG];
  __shared__ int item;
  const int home = blockIdx.x & 7;
  for (int qi = 0; qi < 8; ++qi) {
    const int q = (home + qi) & 7;
    const int base = qstart[q];
    const int len = qstart[q + 1] - base;
    int last = -1;
    for (;;) {
      if (threadIdx.x == 0) item = int(claim_item(&counters[q], len, probe_next(qi, len, last)));
      __syncthreads();
      const int it = __builtin_amdgcn_readfirstlane(item);  // uniform: the job lives in scalar registers
      __syncthreads();  // every thread has read `item` before it is rewritten
      if (it >= len) break;
      last = it;
      const int2 e = items[base + it];
      // products whose rows fit one segment take the paired path (rows
      // copied ahead, schedules in registers), the others stage segments
      const RegJob& jb = jobs[e.x];
      if (jb.m <= SEG && jb.p <= SEG)
        grp_unit<LOG_G, true>(jb, e.y, S, Rb);
      else
        grp_unit<LOG_G, false>(jb, e.y, S, Rb);
      __syncthreads();
    }
  }
}

template <int LOG_G>
int slots_t() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_compress_fwht_grp<LOG_G, true>, 64 << LOG_G, 0) !=
          hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return 512;
  return cus * per_cu;
}

template <int LOG_G>
hipError_t launch_t(const RegJob& j, hipStream_t s) {
  const int lists = j.d << j.lbs;
  k_grp_lists<LOG_G><<<dim3(unsigned(2 * lists)), dim3(64 * LW), 0, s>>>(j);
  const int units = j.nchunks * (lists >> LOG_G);
  const int grid = (units + 7) / 8 * 8;
  if (j.m <= SEG && j.p <= SEG)
    k_compress_fwht_grp<LOG_G, true><<<dim3(unsigned(grid)), dim3(64 << LOG_G), 0, s>>>(j, units);
  else
    k_compress_fwht_grp<LOG_G, false><<<dim3(unsigned(grid)), dim3(64 << LOG_G), 0, s>>>(j, units);
  return hipGetLastError();
}

}  // namespace

int fwht_grp_log_g(int lb) { return lb >= 13 ? 2 : (lb == 12 ? 1 : 0); }

int fwht_grp_supported(int lb, int m, int p) {
  return lb >= 12 && lb <= 30 && m <= MAXSEG * SEG && p <= MAXSEG * SEG;
}

int fwht_grp_slots(int lb) { return fwht_grp_log_g(lb) == 2 ? slots_t<2>() : slots_t<1>(); }

size_t fwht_grp_lists_bytes(int d, int lb, int m, int p) {
  const size_t lists = size_t(d) << reg_lbs(lb);
  return lists * size_t(list_cap(m) + list_cap(p)) * 4 + 2 * lists * MAXSEG * RCAP * sizeof(int) + 256;
}

hipError_t launch_fwht_grp(const RegJob& j, hipStream_t s) {
  if (j.nchunks == 0) return hipSuccess;
  return fwht_grp_log_g(j.lb) == 2 ? launch_t<2>(j, s) : launch_t<1>(j, s);
}

hipError_t launch_fwht_grp_lists_batch(int log_g, const RegJob* jobs, const int2* blocks, int nblocks,
                                          hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  if (log_g == 2)
    k_grp_lists_batch<2><<<dim3(unsigned(nblocks)), dim3(64 * LW), 0, s>>>(jobs, blocks);
  else
    k_grp_lists_batch<1><<<dim3(unsigned(nblocks)), dim3(64 * LW), 0, s>>>(jobs, blocks);
  return hipGetLastError();
}

hipError_t launch_fwht_grp_batch(int log_g, const RegJob* jobs, const int2* items, const int* qstart,
                                   unsigned* counters, int grid, hipStream_t s) {
  if (grid <= 0) return hipSuccess;
  if (log_g == 2)
    k_compress_fwht_grp_batch<2><<<dim3(unsigned(grid)), dim3(256), 0, s>>>(jobs, items, qstart, counters);
  else
    k_compress_fwht_grp_batch<1><<<dim3(unsigned(grid)), dim3(128), 0, s>>>(jobs, items, qstart, counters);
  return hipGetLastError();
}

}  // namespace cm
