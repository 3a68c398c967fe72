// Sliced fused compress, spectral reduction/inverse and cross-slice combine.
//
// Algorithm (reference: matmul_fft.hpp:57-122, matmul_fwht.hpp:53-108):
//   p_t = (1/b) T^-1( sum_k T(S_t a_k) . T(S'_t b_k) )
// where a_k is row k of AT (column k of A), b_k is row k of B, S_t/S'_t the
// count-sketch with (h1_t, s1_t) / (h2_t, s2_t), and T the real DFT (cyclic
// convolution) or the natural-order Walsh-Hadamard transform (XOR
// convolution).  Differences from the reference's loop structure:
//
// * Slicing.  The spectrum of length b = b1*b2 is split into b1 slices of b2
//   frequencies.  A workgroup owns (t, slice, chunk of k).  Its slice of the
//   transform is obtained from a "modulated" count-sketch into b2 buckets
//   followed by a length-b2 transform:
//     FWHT  u[h mod b2] += (-1)^popc((h >> log2 b2) & slice) s(i) a_i,
//           X[slice*b2 + f] = WHT_b2(u)[f]
//     FFT   u[h mod b2] += W_b^(h*slice) s(i) a_i,
//           X[slice + b1*f] = DFT_b2(u)[f]
//   so a product with b up to 2^30 runs through on-chip transforms of at
//   most 8192 points, and the sparse sketch (<= m nonzeros) replaces the
//   first log2(b1) butterfly stages.
// * Two real transforms per complex FFT.  z = S a + i S' b has spectrum
//   Z = A + iB, and A[f]B[f] = (Z[f]^2 - conj(Z[-f])^2)/(4i).  Summed over k
//   this is linear in Q[f] = sum_k Z_k[f]^2, so only Q is accumulated and the
//   inverse reduces to p[h] = Im(IDFT_b(Q)[h]) / (2b).  Rows a_k and b_k are
//   scaled by the power-of-two pair (2^e_k, 2^-e_k) of balance.hip first, so
//   the cancellation costs no accuracy when |a_k| and |b_k| differ.
// * The spectral accumulator lives in registers of the owning workgroup for
//   its whole chunk of k; chunks are summed afterwards in a fixed order
//   (reduce kernel), replacing the reference's per-t omp locks.
#include "device_common.hpp"
#include "kernels.hpp"
#include "lds_transform.hpp"

namespace cm {

namespace {

constexpr int NT = 256;

template <typename T>
struct LdsCfg {
  static constexpr int LB2MAX = sizeof(T) == 4 ? 13 : 12;  // 64 KB of LDS
  static constexpr int MAXR = (1 << LB2MAX) / NT;
};

// sum of the per-chunk partial spectra, written over chunk 0 (each thread
// owns one element and reads every chunk in order: deterministic)
template <typename T, typename ACC>
__device__ __forceinline__ void chunk_sum_body(T* __restrict__ P, int64_t elems, int nchunks, int blk) {
  // chunks are added in order 0, 1, 2, ... (bit-reproducible); the loads of
  // eight chunks are issued before their additions
  constexpr int U = 8;
  if (sizeof(T) == 4 && (elems & 3) == 0) {  // four elements per thread
    const int64_t idx = (int64_t(blk) * NT + threadIdx.x) * 4;
    if (idx >= elems) return;
    ACC s[4] = {ACC(0), ACC(0), ACC(0), ACC(0)};
    int c = 0;
    for (; c + U <= nchunks; c += U) {
      float4 v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = *reinterpret_cast<const float4*>(P + (c + u) * elems + idx);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        s[0] += ACC(v[u].x);
        s[1] += ACC(v[u].y);
        s[2] += ACC(v[u].z);
        s[3] += ACC(v[u].w);
      }
    }
    for (; c < nchunks; ++c) {
      const float4 v = *reinterpret_cast<const float4*>(P + c * elems + idx);
      s[0] += ACC(v.x);
      s[1] += ACC(v.y);
      s[2] += ACC(v.z);
      s[3] += ACC(v.w);
    }
    *reinterpret_cast<float4*>(P + idx) = make_float4(float(s[0]), float(s[1]), float(s[2]), float(s[3]));
    return;
  }
  const int64_t idx = int64_t(blk) * NT + threadIdx.x;
  if (idx >= elems) return;
  ACC s = ACC(0);
  int c = 0;
  for (; c + U <= nchunks; c += U) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = P[(c + u) * elems + idx];
#pragma unroll
    for (int u = 0; u < U; ++u) s += ACC(v[u]);
  }
  for (; c < nchunks; ++c) s += ACC(P[c * elems + idx]);
  P[idx] = T(s);
}

// blocks of the chunk sum: four elements per thread where the layout allows
template <typename T>
__host__ __device__ __forceinline__ int64_t chunk_sum_threads(int64_t elems) {
  return sizeof(T) == 4 && (elems & 3) == 0 ? elems / 4 : elems;
}

template <typename T, typename ACC>
__global__ void __launch_bounds__(NT) k_chunk_sum(T* __restrict__ P, int64_t elems, int nchunks) {
  chunk_sum_body<T, ACC>(P, elems, nchunks, blockIdx.x);
}

// The last radix-R pass of the inverse in registers, each group's outputs
// passed to out(position, value) (stride 1: group g holds g R + j)
template <int R, typename T, bool FWHT, typename F>
__device__ __forceinline__ void reduce_last(ldsx::Elem<T, FWHT>* z, int b2, F&& out) {
  for (int g = threadIdx.x; g < b2 / R; g += NT) {
    ldsx::Elem<T, FWHT> v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = z[ldsx::pad(g * R + j)];
    ldsx::last_pass<R, T, FWHT>(v);
#pragma unroll
    for (int j = 0; j < R; ++j) out(g * R + j, v[j]);
  }
}

template <int RM, typename T, bool FWHT, typename F>
__device__ __forceinline__ void reduce_last_any(ldsx::Elem<T, FWHT>* z, int lb2, F&& out) {
  switch (ldsx::last_radix_log2(lb2, RM)) {
    case 4: reduce_last<16, T, FWHT>(z, 1 << lb2, out); break;
    case 3: reduce_last<8, T, FWHT>(z, 1 << lb2, out); break;
    case 2: reduce_last<4, T, FWHT>(z, 1 << lb2, out); break;
    case 1: reduce_last<2, T, FWHT>(z, 1 << lb2, out); break;
    default: reduce_last<1, T, FWHT>(z, 1 << lb2, out); break;
  }
}

// Sum of the chunks of one (t, slice) spectrum, then its inverse transform
// over the slice in LDS with the radix-16 passes of lds_transform.hpp, the
// last one in registers and written straight out (two LDS passes and three
// barriers for 4096 points, where radix-2 stages took twelve): the
// natural-order WHT, and for the FFT the inverse DFT as conj(DFT(conj Q)),
// the conjugations being exact sign flips.  The slice is written in natural
// order; with one slice and a.P set, the sketch is written instead, as the
// combine kernel computes it (0 + v / b for FWHT, 0 + Im(V conj(W^0)) / 2b
// for FFT, with the same operations so the bits agree).
template <typename T, typename ACC, bool FWHT>
__device__ __forceinline__ void reduce_body(const ReduceArgs& a, int blk) {
  using ldsx::pad;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b2 = 1 << a.lb2;
  const int lbs = a.lb - a.lb2;
  const int b1 = 1 << lbs;
  const int slice = blk & (b1 - 1);
  const int t = blk >> lbs;
  const int tid = threadIdx.x;
  const int64_t chunk_stride = int64_t(a.d) * b1 * b2;
  const int64_t off = (int64_t(t) * b1 + slice) * b2;
  // with the chunks already summed (nchunks 1), the loads of eight elements
  // are issued before their stores; the general loop sums chunk by chunk
  constexpr int U = 8;
  if constexpr (FWHT) {
    T* s = reinterpret_cast<T*>(smem_raw);
    const T* __restrict__ P = static_cast<const T*>(a.partial);
    if (a.nchunks == 1) {
      for (int e0 = tid; e0 < b2; e0 += U * NT) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = e0 + u * NT < b2 ? P[off + e0 + u * NT] : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (e0 + u * NT < b2) s[pad(e0 + u * NT)] = v[u];
      }
    } else {
      for (int e = tid; e < b2; e += NT) {
        ACC sum = ACC(0);
        for (int c = 0; c < a.nchunks; ++c) sum += ACC(P[c * chunk_stride + off + e]);
        s[pad(e)] = T(sum);
      }
    }
    ldsx::lds_transform_head<NT, T, true>(s, a.lb2, 1, 0, nullptr, nullptr);
    if (a.P && b1 == 1) {
      T* P = static_cast<T*>(a.P) + off;
      const T inv = T(1) / T(b2);
      reduce_last_any<ldsx::rmax<T, true>(), T, true>(s, a.lb2, [&](int e, T v) {
        const T sum = T(0) + v;
        P[e] = T(0) + sum * inv;
      });
    } else {
      T* V = static_cast<T*>(a.V) + off;
      reduce_last_any<ldsx::rmax<T, true>(), T, true>(s, a.lb2, [&](int e, T v) { V[e] = v; });
    }
  } else {
    constexpr int RM = 4;  // radix 16
    cpx<T>* z = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* lo = ldsx::twiddle_lds<T, false>(smem_raw, a.lb2);
    cpx<T>* hi = lo + ldsx::TW_LO;
    const cpx<T>* __restrict__ P = static_cast<const cpx<T>*>(a.partial);
    ldsx::fill_twiddles<NT, T>(lo, hi, a.lb2, static_cast<const cpx<T>*>(a.tw), a.lb);
    if (a.nchunks == 1) {
      for (int e0 = tid; e0 < b2; e0 += U * NT) {
        cpx<T> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = e0 + u * NT < b2 ? P[off + e0 + u * NT] : cpx<T>{T(0), T(0)};
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (e0 + u * NT < b2) z[pad(e0 + u * NT)] = cpx<T>{v[u].x, -v[u].y};
      }
    } else {
      for (int e = tid; e < b2; e += NT) {
        ACC sx = ACC(0), sy = ACC(0);
        for (int c = 0; c < a.nchunks; ++c) {
          const cpx<T> v = P[c * chunk_stride + off + e];
          sx += ACC(v.x);
          sy += ACC(v.y);
        }
        z[pad(e)] = {T(sx), -T(sy)};
      }
    }
    ldsx::lds_transform_head<NT, T, false, RM>(z, a.lb2, 1, 0, lo, hi);
    if (a.P && b1 == 1) {
      T* P = static_cast<T*>(a.P) + off;
      const cpx<T> w = static_cast<const cpx<T>*>(a.tw)[0];  // W^0, as the combine reads it
      const T inv = T(1) / T(2 * b2);
      reduce_last_any<RM, T, false>(z, a.lb2, [&](int e, cpx<T> v) {
        const cpx<T> c{v.x, -v.y};
        T im = T(0);
        im += c.y * w.x - c.x * w.y;
        P[ldsx::dif_nat(e, a.lb2, RM)] = T(0) + im * inv;
      });
    } else {
      cpx<T>* V = static_cast<cpx<T>*>(a.V) + off;
      reduce_last_any<RM, T, false>(z, a.lb2, [&](int e, cpx<T> v) {
        V[ldsx::dif_nat(e, a.lb2, RM)] = cpx<T>{v.x, -v.y};
      });
    }
  }
}

template <typename T, typename ACC, bool FWHT>
__global__ void __launch_bounds__(NT) k_reduce_inverse(ReduceArgs a) {
  reduce_body<T, ACC, FWHT>(a, blockIdx.x);
}

// p[t][h] from the per-slice inverse transforms V[t][slice][h mod b2].  A
// zero-padded product (lb_out < lb) sums the 2^(lb - lb_out) positions
// h = ho + f 2^lb_out: the wrapped part of the linear convolution for FFT,
// exact zeros for FWHT.
template <typename T, bool FWHT>
__device__ __forceinline__ void combine_body(const CombineArgs& a, int64_t idx) {
  const int64_t bo = int64_t(1) << a.lb_out;
  if (idx >= int64_t(a.d) * bo) return;
  const int64_t b = int64_t(1) << a.lb;
  const int t = int(idx >> a.lb_out);
  const uint32_t ho = uint32_t(idx & (bo - 1));
  const int b2 = 1 << a.lb2;
  const int b1 = 1 << (a.lb - a.lb2);
  T* P = static_cast<T*>(a.P);
  T total = T(0);
  for (int64_t f = 0; f < (b >> a.lb_out); ++f) {
    const uint32_t h = ho + uint32_t(f << a.lb_out);
    const uint32_t lo = h & uint32_t(b2 - 1), hi = h >> a.lb2;
    // the loads of eight slices are issued before their additions (in
    // slice order, as before)
    constexpr int U = 8;
    if constexpr (FWHT) {
      const T* V = static_cast<const T*>(a.V) + int64_t(t) * b;
      T s = T(0);
      for (int sl0 = 0; sl0 < b1; sl0 += U) {
        T v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = sl0 + u < b1 ? V[int64_t(sl0 + u) * b2 + lo] : T(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (sl0 + u < b1) s += (__popc(uint32_t(sl0 + u) & hi) & 1) ? -v[u] : v[u];
      }
      total += s * (T(1) / T(b));
    } else {
      const cpx<T>* V = static_cast<const cpx<T>*>(a.V) + int64_t(t) * b;
      const cpx<T>* tw = static_cast<const cpx<T>*>(a.tw);
      T im = T(0);
      for (int sl0 = 0; sl0 < b1; sl0 += U) {
        cpx<T> v[U], w[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const bool ok = sl0 + u < b1;
          v[u] = ok ? V[int64_t(sl0 + u) * b2 + lo] : cpx<T>{T(0), T(0)};
          w[u] = ok ? tw[(uint64_t(sl0 + u) * h) & uint64_t(b - 1)] : cpx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (sl0 + u < b1) im += v[u].y * w[u].x - v[u].x * w[u].y;  // Im(v * conj(w))
      }
      total += im * (T(1) / T(2 * b));
    }
  }
  P[idx] = total;
}

template <typename T, bool FWHT>
__global__ void __launch_bounds__(NT) k_combine(CombineArgs a) {
  combine_body<T, FWHT>(a, int64_t(blockIdx.x) * NT + threadIdx.x);
}

// ---------------------------------------------------------------- batches
// FP32 jobs of either variant; the variant and the reduction width are read
// per workgroup.
__global__ void __launch_bounds__(NT) k_chunk_sum_batch(const FinishJob* __restrict__ jobs,
                                                        const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  const ReduceArgs& r = jobs[e.x].r;
  const int64_t elems = (int64_t(r.d) << r.lb) * (r.fwht ? 1 : 2);
  float* P = static_cast<float*>(const_cast<void*>(r.partial));
  if (r.f64_reduce)
    chunk_sum_body<float, double>(P, elems, r.nchunks, e.y);
  else
    chunk_sum_body<float, float>(P, elems, r.nchunks, e.y);
}

// the partials are already summed over chunks (chunk 0 holds the sum; no
// chunk at all for an empty inner dimension)
__global__ void __launch_bounds__(NT) k_reduce_batch(const FinishJob* __restrict__ jobs,
                                                     const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  ReduceArgs r = jobs[e.x].r;
  r.nchunks = r.nchunks > 0 ? 1 : 0;
  if (r.fwht)
    reduce_body<float, float, true>(r, e.y);
  else
    reduce_body<float, float, false>(r, e.y);
}

__global__ void __launch_bounds__(NT) k_combine_batch(const FinishJob* __restrict__ jobs,
                                                      const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  const CombineArgs& c = jobs[e.x].c;
  const int64_t idx = int64_t(e.y) * NT + threadIdx.x;
  if (c.fwht)
    combine_body<float, true>(c, idx);
  else
    combine_body<float, false>(c, idx);
}

// General fused compress (FP64, and FP32 sketches below the register
// kernels' range): one workgroup per (chunk of k, t, slice) keeps its slice
// of the spectral accumulator in registers for its chunk of k.  Each slice
// transform runs as radix-4/8/16 register passes over a padded LDS tile
// (lds_transform.hpp: a barrier per pass, twiddles from an LDS table).  The scatter adds
// with LDS atomics; the accumulator stays in registers, at the transform's
// output positions, and is written in natural frequency order.
// 256 threads at two waves per SIMD: 256 VGPRs a lane hold the accumulator
// (b2 / 256 values) beside radix-16 passes, so a 4096-point FP64 slice takes
// three LDS passes; two workgroups (LDS tiles) share a CU
constexpr int LT = 256;
constexpr int KU = 8;  // scatter keys per thread with loads in flight together
constexpr int KH = 4;  // keys per thread held in registers across k (m + p <= KH LT)
template <typename T>
constexpr int fft_rm() { return 4; }

// log2 of the radix of the transform passes (the last pass may be narrower)
template <typename T, bool FWHT>
constexpr int lds2_rm() { return FWHT ? ldsx::rmax<T, true>() : fft_rm<T>(); }

template <typename T, bool FWHT>
__host__ __device__ __forceinline__ size_t lds2_bytes(int lb2) {
  return FWHT ? ldsx::tile_bytes<T, true>(lb2 + 1) : ldsx::transform_bytes<T, false>(lb2);
}

// RL: log2 of the radix of the last transform pass, last_radix_log2(lb2, RM)
// (one instantiation each: a runtime choice among them costs registers)
template <typename T, bool FWHT, int RL>
__global__ void __launch_bounds__(LT) __attribute__((amdgpu_waves_per_eu(2, 2))) k_compress_lds2(CompressArgs a) {
  using namespace ldsx;
  using E = Elem<T, FWHT>;
  constexpr int PER = (1 << LdsCfg<T>::LB2MAX) / LT;  // accumulator slots per thread
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lb2 = a.lb2, b2 = 1 << lb2;
  const int lbs = a.lb - lb2;
  const int b1 = 1 << lbs;
  int g = blockIdx.x;
  const int slice = g & (b1 - 1);
  g >>= lbs;
  const int t = g % a.d;
  const int chunk = g / a.d;
  const int k0 = chunk * a.kper;
  const int k1 = min(a.n, k0 + a.kper);
  const int tid = threadIdx.x;
  const int keys = a.m + a.p;  // keys [0, m) of A then [0, p) of B
  const T* __restrict__ AT = static_cast<const T*>(a.AT);
  const T* __restrict__ B = static_cast<const T*>(a.B);
  const uint32_t* __restrict__ hA = a.hA + int64_t(t) * a.m;
  const uint32_t* __restrict__ hB = a.hB + int64_t(t) * a.p;
  const int64_t out_off = ((int64_t(chunk) * a.d + t) * b1 + slice) * b2;
  // hash entry and row-k value of a key below `keys`, each one load from a
  // selected address (a select of two loads compiles to a branch per load)
  auto key_entry = [&](int key) -> uint32_t {
    const uint32_t* src = key < a.m ? hA + key : hB + (key - a.m);
    return *src;
  };
  auto row_value = [&](int k, int key) -> T {
    const T* src = key < a.m ? AT + int64_t(k) * a.ld_at + key : B + int64_t(k) * a.ld_b + (key - a.m);
    return *src;
  };
  // When all keys fit KH per thread (uniform), their hash entries stay in
  // registers for the whole chunk and the row values of k + 1 are loaded
  // while k is scattered and transformed.  Key u LT + tid is read at a
  // clamped index, so every load is in bounds.
  const int nu = (keys + LT - 1) / LT;
  const bool held = nu <= KH;
  uint32_t eh[KH];
  T vn[KH];
#pragma unroll
  for (int u = 0; u < KH; ++u) {
    eh[u] = 0u;
    vn[u] = T(0);
    if (held && u < nu) eh[u] = key_entry(min(u * LT + tid, keys - 1));
  }
  auto load_row = [&](int k) {
#pragma unroll
    for (int u = 0; u < KH; ++u)
      if (u < nu) vn[u] = row_value(k, min(u * LT + tid, keys - 1));
  };
  E acc[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) acc[r] = E{};
  // The tile is cleared once; afterwards each k's accumulation clears the
  // elements it reads (the scatter and the transform touch no padding slot).
  if constexpr (FWHT) {
    // A's sketch at elements [0, b2), B's at [b2, 2 b2), one padded tile
    T* sa = reinterpret_cast<T*>(smem_raw);
    constexpr int R = 1 << RL;  // the last pass runs in registers on R elements per group
    const int wwords = int(tile_bytes<T, true>(lb2 + 1) / 16);
    for (int w = tid; w < wwords; w += LT) reinterpret_cast<float4*>(smem_raw)[w] = make_float4(0.f, 0.f, 0.f, 0.f);
    auto scatter = [&](const uint32_t(&e)[KU], const T(&v)[KU], int base) {
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        const int key = base + u * LT + tid;
        if (key < keys) {
          const uint32_t h = hbucket(e[u]);
          const uint32_t neg = hneg(e[u]) ^ (__popc((h >> lb2) & slice) & 1);
          atomicAdd(&sa[pad((key < a.m ? 0 : b2) + int(h & (b2 - 1)))], apply_sign(v[u], neg));
        }
      }
    };
    if (held && k0 < k1) load_row(k0);
    for (int k = k0; k < k1; ++k) {
      __syncthreads();  // the previous k's products are read and their elements cleared
      // KU keys per thread with all their loads in flight before the LDS adds
      // (held keys: one round, from registers)
      for (int base = 0; base < keys; base += LT * KU) {
        uint32_t e[KU];
        T v[KU];
        if (held) {
#pragma unroll
          for (int u = 0; u < KU; ++u) {
            e[u] = u < KH ? eh[u] : 0u;
            v[u] = u < KH ? vn[u] : T(0);
          }
          if (k + 1 < k1) load_row(k + 1);
        } else {
#pragma unroll
          for (int u = 0; u < KU; ++u) {
            const int key = base + u * LT + tid;
            e[u] = 0u;
            v[u] = T(0);
            if (key < keys) {
              e[u] = key_entry(key);
              v[u] = row_value(k, key);
            }
          }
        }
        scatter(e, v, base);
      }
      lds_transform_head<LT, T, true>(sa, lb2, 2, b2, nullptr, nullptr);  // both sketches
#pragma unroll
      for (int i = 0; i < PER / R; ++i) {  // group g = tid + i LT holds elements g R + j
        const int g = tid + i * LT;
        if (g < (b2 >> RL)) {
          T va[R], vb[R];
#pragma unroll
          for (int j = 0; j < R; ++j) {
            va[j] = sa[pad(g * R + j)];
            vb[j] = sa[pad(b2 + g * R + j)];
          }
          last_pass<R, T, true>(va);
          last_pass<R, T, true>(vb);
#pragma unroll
          for (int j = 0; j < R; ++j) {
            acc[i * R + j] += va[j] * vb[j];
            sa[pad(g * R + j)] = T(0);
            sa[pad(b2 + g * R + j)] = T(0);
          }
        }
      }
    }
    T* out = static_cast<T*>(a.partial) + out_off;
#pragma unroll
    for (int i = 0; i < PER / R; ++i) {
      const int g = tid + i * LT;
      if (g < (b2 >> RL)) {
#pragma unroll
        for (int j = 0; j < R; ++j) out[g * R + j] = acc[i * R + j];
      }
    }
  } else {
    constexpr int FRM = fft_rm<T>();
    constexpr int R = 1 << RL;  // the last pass runs in registers on R elements per group
    cpx<T>* z = reinterpret_cast<cpx<T>*>(smem_raw);
    cpx<T>* lo = twiddle_lds<T, false>(smem_raw, lb2);
    cpx<T>* hi = lo + TW_LO;
    const cpx<T>* __restrict__ tw = static_cast<const cpx<T>*>(a.tw);
    fill_twiddles<LT, T>(lo, hi, lb2, tw, a.lb);
    for (int w = tid; w < int(tile_bytes<T, false>(lb2) / 16); w += LT)
      reinterpret_cast<float4*>(smem_raw)[w] = make_float4(0.f, 0.f, 0.f, 0.f);
    const uint64_t bmask = (uint64_t(1) << a.lb) - 1;
    const cpx<T>* __restrict__ bal = static_cast<const cpx<T>*>(a.scale);
    // v: balanced values (power-of-two scales of row k, balance.hip)
    auto scatter = [&](const uint32_t(&e)[KU], const T(&v)[KU], int base) {
      if (slice == 0) {  // modulation W_b^0 = 1: A adds to Re, B to Im only
#pragma unroll
        for (int u = 0; u < KU; ++u) {
          const int key = base + u * LT + tid;
          if (key < keys) {
            cpx<T>* dst = &z[pad(hbucket(e[u]) & (b2 - 1))];
            atomicAdd(key < a.m ? &dst->x : &dst->y, apply_sign(v[u], hneg(e[u])));
          }
        }
      } else {
        cpx<T> w[KU];
#pragma unroll
        for (int u = 0; u < KU; ++u) w[u] = tw[(uint64_t(hbucket(e[u])) * uint64_t(slice)) & bmask];
#pragma unroll
        for (int u = 0; u < KU; ++u) {
          const int key = base + u * LT + tid;
          if (key < keys) {
            const T x = apply_sign(v[u], hneg(e[u]));
            cpx<T>* dst = &z[pad(hbucket(e[u]) & (b2 - 1))];
            if (key < a.m) {
              atomicAdd(&dst->x, x * w[u].x);
              atomicAdd(&dst->y, x * w[u].y);
            } else {  // i x w
              atomicAdd(&dst->x, -x * w[u].y);
              atomicAdd(&dst->y, x * w[u].x);
            }
          }
        }
      }
    };
    if (held && k0 < k1) load_row(k0);
    for (int k = k0; k < k1; ++k) {
      __syncthreads();  // the previous k's squares are read and their elements cleared
      const cpx<T> bk = bal[k];
      // KU keys per thread with all their loads in flight before the LDS adds
      // (held keys: one round, from registers)
      for (int base = 0; base < keys; base += LT * KU) {
        uint32_t e[KU];
        T v[KU];
        if (held) {
#pragma unroll
          for (int u = 0; u < KU; ++u) {
            e[u] = u < KH ? eh[u] : 0u;
            v[u] = u < KH ? vn[u] : T(0);
          }
          if (k + 1 < k1) load_row(k + 1);
        } else {
#pragma unroll
          for (int u = 0; u < KU; ++u) {
            const int key = base + u * LT + tid;
            e[u] = 0u;
            v[u] = T(0);
            if (key < keys) {
              e[u] = key_entry(key);
              v[u] = row_value(k, key);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < KU; ++u) v[u] *= base + u * LT + tid < a.m ? bk.x : bk.y;
        scatter(e, v, base);
      }
      lds_transform_head<LT, T, false, FRM>(z, lb2, 1, 0, lo, hi);
#pragma unroll
      for (int i = 0; i < PER / R; ++i) {  // group g = tid + i LT holds elements g R + j
        const int g = tid + i * LT;
        if (g < (b2 >> RL)) {
          cpx<T> v[R];
#pragma unroll
          for (int j = 0; j < R; ++j) v[j] = z[pad(g * R + j)];
          last_pass<R, T, false>(v);
#pragma unroll
          for (int j = 0; j < R; ++j) {
            acc[i * R + j] = cadd(acc[i * R + j], csqr(v[j]));
            z[pad(g * R + j)] = cpx<T>{T(0), T(0)};
          }
        }
      }
    }
    cpx<T>* out = static_cast<cpx<T>*>(a.partial) + out_off;
#pragma unroll
    for (int i = 0; i < PER / R; ++i) {
      const int g = tid + i * LT;
      if (g < (b2 >> RL)) {
#pragma unroll
        for (int j = 0; j < R; ++j) out[dif_nat(g * R + j, lb2, FRM)] = acc[i * R + j];
      }
    }
  }
}

// the instantiation of the general kernel for slices of 2^lb2
using Lds2Fn = void (*)(CompressArgs);
template <typename T, bool FWHT>
Lds2Fn lds2_kernel(int lb2) {
  switch (ldsx::last_radix_log2(lb2, lds2_rm<T, FWHT>())) {
    case 4: return &k_compress_lds2<T, FWHT, 4>;
    case 3: return &k_compress_lds2<T, FWHT, 3>;
    case 2: return &k_compress_lds2<T, FWHT, 2>;
    case 1: return &k_compress_lds2<T, FWHT, 1>;
    default: return &k_compress_lds2<T, FWHT, 0>;
  }
}

template <typename T, bool FWHT>
hipError_t compress_lds_t(const CompressArgs& a, hipStream_t s) {
  const int b1 = 1 << (a.lb - a.lb2);
  const int64_t grid = int64_t(a.nchunks) * a.d * b1;
  const size_t lds = lds2_bytes<T, FWHT>(a.lb2);
  const Lds2Fn fn = lds2_kernel<T, FWHT>(a.lb2);
  if (lds > 65536 &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                          int(lds)) != hipSuccess)
    return hipErrorInvalidValue;
  fn<<<dim3(unsigned(grid)), dim3(LT), lds, s>>>(a);
  return hipGetLastError();
}

template <typename T, bool FWHT>
hipError_t reduce_t(const ReduceArgs& a0, hipStream_t s) {
  ReduceArgs a = a0;
  const bool wide = a.f64_reduce || sizeof(T) == 8;
  if (a.nchunks > 1) {
    const int64_t elems = (int64_t(a.d) << a.lb) * (FWHT ? 1 : 2);
    const unsigned blocks = unsigned((chunk_sum_threads<T>(elems) + NT - 1) / NT);
    T* P = static_cast<T*>(const_cast<void*>(a.partial));
    if (wide)
      k_chunk_sum<T, double><<<dim3(blocks), dim3(NT), 0, s>>>(P, elems, a.nchunks);
    else
      k_chunk_sum<T, T><<<dim3(blocks), dim3(NT), 0, s>>>(P, elems, a.nchunks);
    a.nchunks = 1;
  }
  const int b1 = 1 << (a.lb - a.lb2);
  const size_t lds = reduce_lds_bytes(FWHT, sizeof(T), a.lb2);
  const unsigned grid = unsigned(a.d * b1);
  // slices of 64 KB need a little more than the default 64 KB of LDS
  const void* fn = wide ? reinterpret_cast<const void*>(&k_reduce_inverse<T, double, FWHT>)
                        : reinterpret_cast<const void*>(&k_reduce_inverse<T, T, FWHT>);
  if (lds > 65536 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) != hipSuccess)
    return hipErrorInvalidValue;
  if (wide)
    k_reduce_inverse<T, double, FWHT><<<dim3(grid), dim3(NT), lds, s>>>(a);
  else
    k_reduce_inverse<T, T, FWHT><<<dim3(grid), dim3(NT), lds, s>>>(a);
  return hipGetLastError();
}

template <typename T, bool FWHT>
hipError_t combine_t(const CombineArgs& a, hipStream_t s) {
  const int64_t total = int64_t(a.d) << a.lb_out;
  k_combine<T, FWHT><<<dim3(unsigned((total + NT - 1) / NT)), dim3(NT), 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace

namespace {
template <typename T, bool FWHT>
int lds2_slots(int lb2) {
  int dev = 0, cus = 0, per_cu = 0;
  const size_t lds = lds2_bytes<T, FWHT>(lb2);
  const Lds2Fn fn = lds2_kernel<T, FWHT>(lb2);
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                              int(lds));
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, LT, lds) != hipSuccess ||
      cus <= 0 || per_cu <= 0)
    return 256;
  return cus * per_cu;
}
}  // namespace

int compress_lds_slots(int fwht, int f64, int lb2) {
  static int cache[2][2][32] = {};  // one occupancy query per configuration
  int& c = cache[fwht != 0][f64 != 0][lb2 & 31];
  if (!c) {
    if (f64)
      c = fwht ? lds2_slots<double, true>(lb2) : lds2_slots<double, false>(lb2);
    else
      c = fwht ? lds2_slots<float, true>(lb2) : lds2_slots<float, false>(lb2);
  }
  return c;
}

int compress_lds_max_lb2(int fwht, int f64) {
  (void)fwht;
  return f64 ? LdsCfg<double>::LB2MAX : LdsCfg<float>::LB2MAX;
}

hipError_t launch_compress_lds(const CompressArgs& a, hipStream_t s) {
  if (a.nchunks == 0) return hipSuccess;
  if (a.f64) return a.fwht ? compress_lds_t<double, true>(a, s) : compress_lds_t<double, false>(a, s);
  return a.fwht ? compress_lds_t<float, true>(a, s) : compress_lds_t<float, false>(a, s);
}

hipError_t launch_reduce_inverse(const ReduceArgs& a, hipStream_t s) {
  if (a.f64) return a.fwht ? reduce_t<double, true>(a, s) : reduce_t<double, false>(a, s);
  return a.fwht ? reduce_t<float, true>(a, s) : reduce_t<float, false>(a, s);
}

int finish_batch_blocks(const FinishJob& j, int* sum_blocks, int* red_blocks) {
  const int64_t elems = (int64_t(j.r.d) << j.r.lb) * (j.r.fwht ? 1 : 2);
  *sum_blocks = j.r.nchunks > 1 ? int((chunk_sum_threads<float>(elems) + NT - 1) / NT) : 0;
  *red_blocks = j.r.d << (j.r.lb - j.r.lb2);
  return int(((int64_t(j.c.d) << j.c.lb_out) + NT - 1) / NT);
}

hipError_t launch_finish_batch(const FinishJob* jobs, const int2* sum_blocks, int n_sum,
                               const int2* red_blocks, int n_red, size_t red_lds,
                               const int2* comb_blocks, int n_comb, hipStream_t s) {
  if (n_sum > 0) k_chunk_sum_batch<<<dim3(unsigned(n_sum)), dim3(NT), 0, s>>>(jobs, sum_blocks);
  if (n_red > 0) k_reduce_batch<<<dim3(unsigned(n_red)), dim3(NT), red_lds, s>>>(jobs, red_blocks);
  if (n_comb > 0) k_combine_batch<<<dim3(unsigned(n_comb)), dim3(NT), 0, s>>>(jobs, comb_blocks);
  return hipGetLastError();
}

hipError_t launch_combine(const CombineArgs& a, hipStream_t s) {
  if (a.f64) return a.fwht ? combine_t<double, true>(a, s) : combine_t<double, false>(a, s);
  return a.fwht ? combine_t<float, true>(a, s) : combine_t<float, false>(a, s);
}

}  // namespace cm
