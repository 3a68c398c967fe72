// Slice-group FP32 FFT compress: one scatter of the pair-packed row for G
// slices (the FFT counterpart of compress_grp.hip).
//
// Slice s of the spectrum X[s + b1 f] (b1 = b / 2048 slices of 2048) is the
// 2048-point DFT of the modulated sketch z_s[l] = sum_{h mod 2048 = l}
// c_h W_b^(h s).  A workgroup of G = 2^LOG_G waves owns the strided slices
// s_j = s0 + j b1/G (j < G) of one repetition.  With h = l + 2048 q and
// q = hw + G hg (hw = q mod G):
//   W_b^(h s_j) = W_b^(h s0) W_G^(hw j) W_{2048 G}^(l j),
// so every slice of the group follows from ONE scatter with the slice-s0
// coefficients c_h W_b^(h s0) into G partitions P[hw][l]:
//   z_{s_j}[l] = W_{2048 G}^(l j) sum_hw W_G^(hw j) P[hw][l],
// a G-point DFT across the partitions and a twiddle per (l, j).  Wave w
// scatters the keys of partition hw = w (conflict-free schedule, as in the
// single-wave kernel), the waves then run the cross-partition step on a
// share of the positions l each, and wave j runs the 2048-point register
// DFT (fft2048_reg.hpp) of slice s_j and accumulates Q += z^2.  The partial
// spectra keep the [chunk][t][slice][2048] layout.  Every value of row k
// enters the scatter times its power-of-two balance scale (balance.hip):
// register-staged segments are scaled as they are staged, the first segment
// (copied straight into LDS) as it is scattered.
#include "device_common.hpp"
#include "fft2048_reg.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

using f2k::R;
using f2k::fq_of;

constexpr int LB2 = 11;
constexpr int B2 = 1 << LB2;
constexpr int SLEN = 2112;   // complex slots per partition (transpose + dummies)
constexpr int SEG = 2048;    // keys staged at a time
constexpr int RCAP = 8;
constexpr int MAXSEG = 64;   // m + p <= MAXSEG * SEG = 131072
constexpr int GB = 8;

__host__ __device__ __forceinline__ int64_t list_cap(int keys) {
  const int64_t segs = (int64_t(keys) + SEG - 1) / SEG;
  return int64_t(keys) + segs * 64 * RCAP;
}

struct Layout {
  float2* C;      // per-item coefficients
  uint32_t* L;    // entries
  int* segcnt;    // [list][MAXSEG][RCAP] group counts
  int64_t cap;
  int lists;      // d * b1, one per (t, s0, partition)
};

__host__ __device__ __forceinline__ Layout layout(const RegJob& j) {
  Layout o;
  o.lists = j.d << j.lbs;
  o.cap = list_cap(j.m + j.p);
  o.C = static_cast<float2*>(j.lists);
  o.L = reinterpret_cast<uint32_t*>(o.C + o.lists * o.cap);
  o.segcnt = reinterpret_cast<int*>(o.L + o.lists * o.cap);
  return o;
}

// Schedule of list g = (t * (b1/G) + s0) * G + w over the concatenated keys
// [0, m + p): the keys of partition hw = w, bucket l = h mod 2048,
// coefficient s W_b^(h s0) (times i for B).  Entry: l << 3 (byte offset of
// a complex, or a dummy slot 2048 + lane) | key offset in the segment << 18.
// LW waves build one list, as in the FWHT slice-group builder: wave bw takes
// keys [bw SEG/LW, (bw+1) SEG/LW) of each segment; ranks and round runs are
// offset by the earlier waves' counts, so the list is the one a single wave
// walking the keys in order would build.
constexpr int LW = 4;

template <int LOG_G>
__device__ __forceinline__ void fft_grp_lists_body(const RegJob& j, int g) {
  constexpr int G = 1 << LOG_G;
  constexpr int KW = SEG / LW;   // keys per wave and segment
  constexpr int Q = KW / 64;     // keys per lane and segment
  __shared__ uint32_t cnt[B2];     // occurrences in earlier segments
  __shared__ uint32_t cw[LW][B2];  // occurrences in this segment, per wave
  __shared__ int nrw[LW][RCAP];    // this segment's keys per wave and round
  const Layout lay = layout(j);
  const int b1 = 1 << j.lbs;
  const int t = g >> j.lbs;
  const int rest = g & (b1 - 1);
  const uint64_t s0 = uint64_t(rest >> LOG_G);
  const uint32_t w = uint32_t(rest & (G - 1));
  const uint64_t bmask = (uint64_t(1) << j.lb) - 1;
  const float2* __restrict__ tw = static_cast<const float2*>(j.tw);
  const int m = j.m, p = j.p, keys = m + p;
  const uint32_t* __restrict__ hA = j.hA + int64_t(t) * m;
  const uint32_t* __restrict__ hB = j.hB + int64_t(t) * p;
  uint32_t* __restrict__ Ls = lay.L + g * lay.cap;
  float2* __restrict__ Cs = lay.C + g * lay.cap;
  const int tid = threadIdx.x;
  const int bw = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  for (int e = tid; e < B2; e += 64 * LW) cnt[e] = 0;
  for (int e = tid; e < LW * B2; e += 64 * LW) (&cw[0][0])[e] = 0;
  int out = 0;
  for (int base0 = 0; base0 < keys; base0 += SEG) {
    const int sl = min(SEG, keys - base0);
    // this wave's hash entries of the segment in flight at once
    uint32_t ek[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int kk = bw * KW + q * 64 + lane;
      const int key = base0 + kk;
      ek[q] = kk < sl ? (key < m ? hA[key] : hB[key - m]) : 0u;
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
    // every key's twiddle W_b^(h s0), all loads in flight at once (inside
    // the per-key branch below they would be issued one after another)
    float2 wq[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) wq[q] = tw[(uint64_t(hbucket(ek[q])) * s0) & bmask];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = rk[q];
      const int kk = bw * KW + q * 64 + lane;
      uint32_t ent = 0;
      float2 coef = make_float2(0.f, 0.f);
      if (r >= 0) {
        const bool isb = base0 + kk >= m;
        const uint32_t h = hbucket(ek[q]);
        const float2 wv = wq[q];
        const float sg = hneg(ek[q]) ? -1.f : 1.f;
        coef = isb ? make_float2(-sg * wv.y, sg * wv.x) : make_float2(sg * wv.x, sg * wv.y);
        ent = ((h & (B2 - 1)) << 3) | (uint32_t(kk) << 18);
      }
      const int64_t pos = round_slot(r, lane, run);
      if (r >= 0) {
        Ls[pos] = ent;
        Cs[pos] = coef;
      }
    }
#pragma unroll
    for (int rr = 0; rr < RCAP; ++rr) {  // pad rounds to whole groups (dummy slots)
      const int full = (tot[rr] + 63) / 64 * 64;
      for (int slot = tot[rr] + tid; slot < full; slot += 64 * LW) {
        const int64_t pos = int64_t(base[rr]) * 64 + slot;
        Ls[pos] = uint32_t(B2 + (slot & 63)) << 3;
        Cs[pos] = make_float2(0.f, 0.f);
      }
    }
    if (tid < RCAP) {
      int all = 0;
#pragma unroll
      for (int w2 = 0; w2 < LW; ++w2) all += nrw[w2][tid];
      const int64_t row = int64_t(g) * MAXSEG + base0 / SEG;
      lay.segcnt[row * RCAP + tid] = (all + 63) / 64;
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
__global__ void __launch_bounds__(64 * LW) k_fft_grp_lists(RegJob j) { fft_grp_lists_body<LOG_G>(j, blockIdx.x); }

template <int LOG_G>
__global__ void __launch_bounds__(64 * LW) k_fft_grp_lists_batch(const RegJob* __restrict__ jobs,
                                                          const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  fft_grp_lists_body<LOG_G>(jobs[e.x], e.y);
}

// Balance scale of a staged value at byte offset `off` of the segment: the
// staged rows of register-staged segments already carry it (SC = 0); a
// segment copied straight into LDS carries raw values, and its keys below
// byte offset `ab` are A's (scale bal.x), the others B's (SC = 1).
template <int SC>
__device__ __forceinline__ float staged_scale(uint32_t off, float2 bal, uint32_t ab) {
  if constexpr (SC == 0) return 1.f;
  return off < ab ? bal.x : bal.y;
}

// KIND 0 plain writes, 1 read-add-writes, 2 LDS atomics (as in the
// single-wave kernel)
template <int KIND, int SC>
__device__ __forceinline__ void scatter_section(const uint32_t* __restrict__ lp,
                                                const float2* __restrict__ cp, int ng, float2* S,
                                                const float* Rb, int lane, float2 bal, uint32_t ab) {
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
    for (int q = 0; q < GB; ++q) {
      const uint32_t o = (ent[q] >> 16) & 0x1ffcu;
      v[q] = *reinterpret_cast<const float*>(Rbb + o);
      if (SC) v[q] *= staged_scale<SC>(o, bal, ab);
    }
#pragma unroll
    for (int q = 0; q < GB; ++q) {
      if (q < nb) {
        const float2 c = make_float2(v[q] * cf[q].x, v[q] * cf[q].y);
        float2* d = reinterpret_cast<float2*>(Sb + (ent[q] & 0xfff8u));
        if (KIND == 2) {
          float* df = reinterpret_cast<float*>(d);
          __hip_atomic_fetch_add(df, c.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(df + 1, c.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else if (KIND == 1) {
          const float2 o = *d;
          *d = make_float2(o.x + c.x, o.y + c.y);
        } else {
          *d = c;
        }
      }
    }
  }
}

// A segment's whole schedule in one batch when it has at most SEGB groups:
// the entries and coefficients of all three sections are loaded together
// (one load latency per segment instead of one per section), then each
// group is written by its kind: plain below n0, read-add-write below
// n0 + n1, atomic after.  The counts are uniform, so the tests are scalar.
constexpr int SEGB = 12;

template <int SC>
__device__ __forceinline__ void scatter_segment(const uint32_t* __restrict__ lp,
                                                const float2* __restrict__ cp, int n0, int n1,
                                                int ng, float2* S, const float* Rb, int lane,
                                                float2 bal, uint32_t ab) {
  const char* Rbb = reinterpret_cast<const char*>(Rb);
  char* Sb = reinterpret_cast<char*>(S);
  uint32_t ent[SEGB];
  float2 cf[SEGB];
#pragma unroll
  for (int g = 0; g < SEGB; ++g) {
    ent[g] = 0u;
    cf[g] = make_float2(0.f, 0.f);
    if (g < ng) {
      ent[g] = lp[g * 64 + lane];
      cf[g] = cp[g * 64 + lane];
    }
  }
  float v[SEGB];
#pragma unroll
  for (int g = 0; g < SEGB; ++g) {
    const uint32_t o = (ent[g] >> 16) & 0x1ffcu;
    v[g] = g < ng ? *reinterpret_cast<const float*>(Rbb + o) : 0.f;
    if (SC) v[g] *= staged_scale<SC>(o, bal, ab);
  }
#pragma unroll
  for (int g = 0; g < SEGB; ++g) {
    if (g >= ng) continue;
    const float2 c = make_float2(v[g] * cf[g].x, v[g] * cf[g].y);
    float2* d = reinterpret_cast<float2*>(Sb + (ent[g] & 0xfff8u));
    if (g < n0) {
      *d = c;
    } else if (g < n0 + n1) {
      const float2 o = *d;
      *d = make_float2(o.x + c.x, o.y + c.y);
    } else {
      float* df = reinterpret_cast<float*>(d);
      __hip_atomic_fetch_add(df, c.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_add(df + 1, c.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
}

// Section counts (plain, read-add-write, atomic groups) of the first
// SEGC segments, read once per unit: they do not depend on k.
constexpr int SEGC = 2;
struct SegCounts {
  int n0[SEGC], n1[SEGC], n2[SEGC];
};

__device__ __forceinline__ void seg_counts(const int* __restrict__ nr, int& n0, int& n1, int& n2) {
  int nw = 0;
  for (int r = 1; r < RCAP - 1; ++r) nw += nr[r];
  n0 = __builtin_amdgcn_readfirstlane(nr[0]);
  n1 = __builtin_amdgcn_readfirstlane(nw);
  n2 = __builtin_amdgcn_readfirstlane(nr[RCAP - 1]);
}

// W_128^k = exp(-2 pi i k / 128) for k < 32, rounded from double
__constant__ float2 kW128[32] = {
    {1.0f, -0.0f}, {0.99879545620517241f, -0.049067674327418015f},
    {0.99518472667219693f, -0.098017140329560604f}, {0.98917650996478101f, -0.14673047445536175f},
    {0.98078528040323043f, -0.19509032201612825f}, {0.97003125319454397f, -0.24298017990326387f},
    {0.95694033573220882f, -0.29028467725446233f}, {0.94154406518302081f, -0.33688985339222005f},
    {0.92387953251128674f, -0.38268343236508978f}, {0.90398929312344334f, -0.42755509343028208f},
    {0.88192126434835505f, -0.47139673682599764f}, {0.85772861000027212f, -0.51410274419322166f},
    {0.83146961230254524f, -0.55557023301960218f}, {0.80320753148064494f, -0.59569930449243336f},
    {0.77301045336273699f, -0.63439328416364549f}, {0.74095112535495911f, -0.67155895484701833f},
    {0.70710678118654757f, -0.70710678118654746f}, {0.67155895484701833f, -0.74095112535495911f},
    {0.63439328416364549f, -0.77301045336273699f}, {0.59569930449243347f, -0.80320753148064483f},
    {0.55557023301960229f, -0.83146961230254524f}, {0.51410274419322166f, -0.85772861000027212f},
    {0.47139673682599781f, -0.88192126434835494f}, {0.4275550934302822f, -0.90398929312344334f},
    {0.38268343236508984f, -0.92387953251128674f}, {0.33688985339222005f, -0.94154406518302081f},
    {0.29028467725446233f, -0.95694033573220894f}, {0.24298017990326398f, -0.97003125319454397f},
    {0.19509032201612833f, -0.98078528040323043f}, {0.14673047445536175f, -0.98917650996478101f},
    {0.09801714032956077f, -0.99518472667219682f}, {0.049067674327418126f, -0.99879545620517241f},
};

// G-point DFT (natural order, W_G = exp(-2 pi i / G)), G in {2, 4}
template <int G>
__device__ __forceinline__ void dft_g(float2 (&y)[G]) {
  if constexpr (G == 2) {
    const float2 a = y[0], b = y[1];
    y[0] = make_float2(a.x + b.x, a.y + b.y);
    y[1] = make_float2(a.x - b.x, a.y - b.y);
  } else {
    const float2 a = y[0], b = y[1], c = y[2], d = y[3];
    const float2 ac = make_float2(a.x + c.x, a.y + c.y), am = make_float2(a.x - c.x, a.y - c.y);
    const float2 bd = make_float2(b.x + d.x, b.y + d.y), bm = make_float2(b.x - d.x, b.y - d.y);
    y[0] = make_float2(ac.x + bd.x, ac.y + bd.y);
    y[2] = make_float2(ac.x - bd.x, ac.y - bd.y);
    y[1] = make_float2(am.x + bm.y, am.y - bm.x);  // am - i bm
    y[3] = make_float2(am.x - bm.y, am.y + bm.x);  // am + i bm
  }
}

// Start copying keys [0, sl) of the concatenated row (A's entries, then
// B's) into LDS dst[0, SEG) with global_load_lds, 4 bytes per lane and no
// registers; lanes past the segment copy its last key.
template <int G>
__device__ __forceinline__ void copy_seg0_async(const float* __restrict__ ar, int m,
                                                const float* __restrict__ br, int sl, float* dst,
                                                int tid) {
  // the wave's base, uniform: the LDS destinations (M0) are scalar
  const int wbase = __builtin_amdgcn_readfirstlane(tid & ~63);
  if (m >= SEG) {  // the whole segment comes from A: no clamp, no select
#pragma unroll
    for (int q = 0; q < SEG / (64 * G); ++q) {
      const int e0 = q * 64 * G + wbase;
      const float* rq = ar + e0;  // uniform; the lane offset is unsigned
      __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(rq + uint32_t(tid & 63)),
                                       (__attribute__((address_space(3))) void*)(dst + e0), 4, 0, 0);
    }
    return;
  }
  // opaque per call, so the per-lane addresses are recomputed here instead
  // of being hoisted out of the k loop and kept live
  int lane_id = tid;
  asm volatile("" : "+v"(lane_id));
#pragma unroll
  for (int q = 0; q < SEG / (64 * G); ++q) {
    const int key = min(q * 64 * G + lane_id, sl - 1);
    const float* src = key < m ? ar + key : br + (key - m);
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)(dst + q * 64 * G + wbase), 4,
                                     0, 0);
  }
}

// unit u = (chunk, t, s0): wave j owns slice s0 + j b1/G
// ASYNC: copy the first segment of row k + 1 during the transforms of k
// (single-product launches; measured slower in the persistent batch kernel,
// whose units are mostly short).
template <int LOG_G, bool ASYNC>
__device__ __forceinline__ void fft_grp_unit(const RegJob& j, int u, float2* S, float* Rb) {
  constexpr int G = 1 << LOG_G;
  const Layout lay = layout(j);
  const int b1 = 1 << j.lbs;
  const int per_chunk = lay.lists >> LOG_G;       // (t, s0) pairs
  const int chunk = u / per_chunk;
  const int ts0 = u - chunk * per_chunk;            // t * (b1/G) + s0
  const int t = ts0 / (b1 >> LOG_G);
  const int s0 = ts0 - t * (b1 >> LOG_G);
  // The wave index and the section counts are uniform.  The batch kernel
  // (register staging) keeps them in scalar registers, so its scatter loops
  // run under scalar control: 2% faster on a 512-product batch.  The same
  // form measured 4% slower in the single-product kernel, which keeps
  // per-lane loop control.
  const int w = ASYNC ? int(threadIdx.x >> 6) : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int tid = threadIdx.x;
  const int slice = s0 + w * (b1 >> LOG_G);
  const int g = ts0 * G + w;                        // this wave's list
  const int k0 = chunk * j.kper;
  const int k1 = min(j.n, k0 + j.kper);
  const uint32_t* __restrict__ ls = lay.L + g * lay.cap;
  const float2* __restrict__ cs = lay.C + g * lay.cap;
  const int* __restrict__ sc = lay.segcnt + int64_t(g) * MAXSEG * RCAP;
  const float2* __restrict__ t2 = static_cast<const float2*>(j.t2);
  const float2* __restrict__ tw = static_cast<const float2*>(j.tw);
  const int keys = j.m + j.p;
  const int nseg = (keys + SEG - 1) / SEG;
  // twiddle W_{2048 G}^(l w) for this wave's output partition = W_b^(l w b/(2048 G))
  const int tw_shift = j.lbs - LOG_G;
  float2* Sw = S + w * SLEN;

  float qr[R], qi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) qr[r] = qi[r] = 0.f;
  // the first segment of row k + 1 is copied straight into LDS while the
  // transforms of k run; later segments are staged through registers, whose
  // loads are issued before the barrier and overlap the other waves'
  // scatter of the previous segment (an LDS copy could only start after it)
  if (ASYNC && k0 < k1 && nseg > 0)
    copy_seg0_async<G>(j.AT + int64_t(k0) * j.ld_at, j.m, j.B + int64_t(k0) * j.ld_b, min(SEG, keys), Rb, tid);
  SegCounts scn;
#pragma unroll
  for (int q = 0; q < SEGC; ++q) {
    scn.n0[q] = scn.n1[q] = scn.n2[q] = 0;
    if (q < nseg) seg_counts(sc + q * RCAP, scn.n0[q], scn.n1[q], scn.n2[q]);
  }
  // A's keys of the first segment end at this byte offset (for the raw
  // values the LDS copy brings)
  const uint32_t ab0 = uint32_t(min(j.m, SEG)) * 4u;
  for (int k = k0; k < k1; ++k) {
    const float* ar = j.AT + int64_t(k) * j.ld_at;
    const float* br = j.B + int64_t(k) * j.ld_b;
    const float2 bal = j.scale[k];  // power-of-two balance of row k (balance.hip)
    {
      float4* S4 = reinterpret_cast<float4*>(Sw);
#pragma unroll
      for (int q = 0; q < B2 / 2 / 64; ++q) S4[q * 64 + lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    int64_t off = 0;
    for (int seg = 0; seg < nseg; ++seg) {
      if (ASYNC && seg == 0) {
        __builtin_amdgcn_s_waitcnt(0);  // this wave's copies of the segment have landed
        __syncthreads();                // and every wave's
      } else {
        const int base0 = seg * SEG;
        const int sl = min(SEG, keys - base0);
        float v[SEG / (64 * G)];
        if (sl == SEG && (base0 + SEG <= j.m || base0 >= j.m)) {
          // a full segment from one operand: a uniform base and the lane's
          // unsigned offset, no clamp and no select
          const float* src = base0 >= j.m ? br + (base0 - j.m) : ar + base0;
          const float sc = base0 >= j.m ? bal.y : bal.x;
#pragma unroll
          for (int q = 0; q < SEG / (64 * G); ++q) v[q] = (src + q * 64 * G)[uint32_t(tid)] * sc;
        } else {
#pragma unroll
          for (int q = 0; q < SEG / (64 * G); ++q) {
            const int key = base0 + min(q * 64 * G + tid, sl - 1);
            const bool from_a = key < j.m;  // one load, from a selected address
            const float* src = from_a ? ar + key : br + (key - j.m);
            v[q] = *src * (from_a ? bal.x : bal.y);
          }
        }
        __syncthreads();  // every wave is done with the previous segment
#pragma unroll
        for (int q = 0; q < SEG / (64 * G); ++q) Rb[q * 64 * G + tid] = v[q];
        __syncthreads();
      }
      int n0, n1, n2;  // uniform: from the per-unit counts (a select chain) or read
      if (seg < SEGC) {
        n0 = scn.n0[0];
        n1 = scn.n1[0];
        n2 = scn.n2[0];
#pragma unroll
        for (int q = 1; q < SEGC; ++q)
          if (seg == q) {
            n0 = scn.n0[q];
            n1 = scn.n1[q];
            n2 = scn.n2[q];
          }
      } else {
        seg_counts(sc + seg * RCAP, n0, n1, n2);
      }
      if (ASYNC && seg == 0) {  // raw values from the LDS copy: scaled here
        if (n0 + n1 + n2 <= SEGB) {
          scatter_segment<1>(ls + off, cs + off, n0, n1, n0 + n1 + n2, Sw, Rb, lane, bal, ab0);
          off += int64_t(n0 + n1 + n2) * 64;
          continue;
        }
        scatter_section<0, 1>(ls + off, cs + off, n0, Sw, Rb, lane, bal, ab0);
        off += int64_t(n0) * 64;
        scatter_section<1, 1>(ls + off, cs + off, n1, Sw, Rb, lane, bal, ab0);
        off += int64_t(n1) * 64;
        scatter_section<2, 1>(ls + off, cs + off, n2, Sw, Rb, lane, bal, ab0);
        off += int64_t(n2) * 64;
        continue;
      }
      if (n0 + n1 + n2 <= SEGB) {  // a short segment schedule in one batch
        scatter_segment<0>(ls + off, cs + off, n0, n1, n0 + n1 + n2, Sw, Rb, lane, bal, 0u);
        off += int64_t(n0 + n1 + n2) * 64;
        continue;
      }
      scatter_section<0, 0>(ls + off, cs + off, n0, Sw, Rb, lane, bal, 0u);
      off += int64_t(n0) * 64;
      scatter_section<1, 0>(ls + off, cs + off, n1, Sw, Rb, lane, bal, 0u);
      off += int64_t(n1) * 64;
      scatter_section<2, 0>(ls + off, cs + off, n2, Sw, Rb, lane, bal, 0u);
      off += int64_t(n2) * 64;
    }
    __syncthreads();  // all partitions complete, and every wave is done reading Rb
    if (ASYNC && nseg > 0 && k + 1 < k1)
      copy_seg0_async<G>(ar + j.ld_at, j.m, br + j.ld_b, min(SEG, keys), Rb, tid);
    {
      // positions l of this wave: [w 2048/G, (w+1) 2048/G); output
      // partition jj gets the DFT over hw times W_{2048 G}^(l jj)
      constexpr int PER = R / G;
      const int l0 = w * (B2 / G);
      // W_{2048 G}^(l jj) for l = l0 + lane + 64 q is the per-lane
      // W_{2048 G}^((l0 + lane) jj), one table load per k, times the
      // constant W_{32 G}^(q jj) = W_128^(q jj 4 / G): no load inside the loop
      float2 twb[G];
#pragma unroll
      for (int jj = 1; jj < G; ++jj) twb[jj] = tw[uint32_t(((l0 + lane) * jj) << tw_shift)];  // index < b
#pragma unroll 1  // unrolled, all positions' loads would be live at once
      for (int q = 0; q < PER; ++q) {
        const int l = l0 + q * 64 + lane;
        float2 y[G];
#pragma unroll
        for (int hw = 0; hw < G; ++hw) y[hw] = S[hw * SLEN + l];
        dft_g<G>(y);
#pragma unroll
        for (int jj = 1; jj < G; ++jj) {
          const float2 c = kW128[(q * jj * (4 / G)) & 31];
          const float2 wv = make_float2(twb[jj].x * c.x - twb[jj].y * c.y, twb[jj].x * c.y + twb[jj].y * c.x);
          y[jj] = make_float2(y[jj].x * wv.x - y[jj].y * wv.y, y[jj].x * wv.y + y[jj].y * wv.x);
        }
#pragma unroll
        for (int jj = 0; jj < G; ++jj) S[jj * SLEN + l] = y[jj];
      }
    }
    __syncthreads();
    float re[R], im[R];
    f2k::dft2048_lds<true>(Sw, t2, lane, re, im);
#pragma unroll
    for (int r = 0; r < R; ++r) {  // Q += z^2
      qr[r] += re[r] * re[r] - im[r] * im[r];
      qi[r] = fmaf(2.f * re[r], im[r], qi[r]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) Sw[fq_of(lane, r)] = make_float2(qr[r], qi[r]);
  __syncthreads();
  float2* out = reinterpret_cast<float2*>(j.partial) +
                (int64_t(chunk) * lay.lists + int64_t(t) * b1 + slice) * B2;
#pragma unroll
  for (int r = 0; r < R; ++r) out[r * 64 + lane] = Sw[r * 64 + lane];
}

// Two waves per SIMD is worth a small spill: measured 2.04 ms against
// 3.34 ms at one wave per SIMD without the spill (fft 2048^3, b = 2^14).
template <int LOG_G>
__global__ void __launch_bounds__(64 << LOG_G) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_compress_fft_grp(RegJob j, int units) {
  __shared__ __align__(16) float2 S[SLEN << LOG_G];
  __shared__ __align__(16) float Rb[SEG];
  const int G = gridDim.x;
  const int g = blockIdx.x;
  const int u = (g & 7) * (G >> 3) + (g >> 3);  // XCD-aware numbering
  if (u >= units) return;
  fft_grp_unit<LOG_G, true>(j, u, S, Rb);
}

// Persistent form for many products: the XCD work queues of the single-wave
// batch kernels; thread 0 claims each item and the workgroup reads it from
// LDS.
template <int LOG_G>
__global__ void __launch_bounds__(64 << LOG_G) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_compress_fft_grp_batch(const RegJob* __restrict__ jobs, const int2* __restrict__ items,
                  const int* __restrict__ qstart, unsigned* __restrict__ counters) {
  __shared__ __align__(16) float2 S[SLEN << LOG_G];
  __shared__ __align__(16) float Rb[SEG];
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
      fft_grp_unit<LOG_G, false>(jobs[e.x], e.y, S, Rb);
      __syncthreads();
    }
  }
}

template <int LOG_G>
int slots_t() {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_compress_fft_grp<LOG_G>, 64 << LOG_G, 0) !=
          hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return 512;
  return cus * per_cu;
}

template <int LOG_G>
hipError_t launch_t(const RegJob& j, hipStream_t s) {
  const int lists = j.d << j.lbs;
  k_fft_grp_lists<LOG_G><<<dim3(unsigned(lists)), dim3(64 * LW), 0, s>>>(j);
  const int units = j.nchunks * (lists >> LOG_G);
  const int grid = (units + 7) / 8 * 8;
  k_compress_fft_grp<LOG_G><<<dim3(unsigned(grid)), dim3(64 << LOG_G), 0, s>>>(j, units);
  return hipGetLastError();
}

}  // namespace

int fft_grp_log_g(int lb) { return lb >= 13 ? 2 : (lb == 12 ? 1 : 0); }

int fft_grp_supported(int lb, int m, int p) {
  return lb >= 12 && lb <= 30 && int64_t(m) + p <= int64_t(MAXSEG) * SEG;
}

int fft_grp_slots(int lb) { return fft_grp_log_g(lb) == 2 ? slots_t<2>() : slots_t<1>(); }

size_t fft_grp_lists_bytes(int d, int lb, int m, int p) {
  const size_t lists = size_t(d) << reg_lbs(lb);
  return lists * size_t(list_cap(m + p)) * 12 + lists * MAXSEG * RCAP * sizeof(int) + 256;
}

hipError_t launch_fft_grp(const RegJob& j, hipStream_t s) {
  if (j.nchunks == 0) return hipSuccess;
  return fft_grp_log_g(j.lb) == 2 ? launch_t<2>(j, s) : launch_t<1>(j, s);
}

hipError_t launch_fft_grp_lists_batch(int log_g, const RegJob* jobs, const int2* blocks, int nblocks,
                                          hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  if (log_g == 2)
    k_fft_grp_lists_batch<2><<<dim3(unsigned(nblocks)), dim3(64 * LW), 0, s>>>(jobs, blocks);
  else
    k_fft_grp_lists_batch<1><<<dim3(unsigned(nblocks)), dim3(64 * LW), 0, s>>>(jobs, blocks);
  return hipGetLastError();
}

hipError_t launch_fft_grp_batch(int log_g, const RegJob* jobs, const int2* items, const int* qstart,
                                   unsigned* counters, int grid, hipStream_t s) {
  if (grid <= 0) return hipSuccess;
  if (log_g == 2)
    k_compress_fft_grp_batch<2><<<dim3(unsigned(grid)), dim3(256), 0, s>>>(jobs, items, qstart, counters);
  else
    k_compress_fft_grp_batch<1><<<dim3(unsigned(grid)), dim3(128), 0, s>>>(jobs, items, qstart, counters);
  return hipGetLastError();
}

}  // namespace cm
