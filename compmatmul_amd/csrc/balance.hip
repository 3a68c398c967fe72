// Per-inner-index balance of the pair-packed FFT.
//
// The FFT paths transform z_k = S a_k + i S' b_k (a_k = row k of AT, b_k =
// row k of B), accumulate Q = sum_k Z_k^2 and extract
//   A_k[f] B_k[f] = (Z_k[f]^2 - conj(Z_k[-f])^2) / 4i
// by cancellation.  What cancels is |A_k|^2 + |B_k|^2, so the rounding of
// the transform of z_k, relative to the product, grows like |A_k|/|B_k| +
// |B_k|/|A_k|.  The reference transforms the two sketches separately
// (matmul_fft.hpp:95-105) and has no such term.
//
// Scaling a_k by 2^e_k and b_k by 2^-e_k leaves every product A_k B_k
// unchanged (power-of-two factors are exact) and, with
//   e_k = round(log2(|b_k| / |a_k|) / 2),
// makes both halves of z_k about equally large, so the error is that of a
// balanced pair at every |a_k| / |b_k|.  The count-sketch keeps the norm of
// a row in expectation, so the row norms stand for the sketch spectra.
// A row pair with |a_k| = 0 or |b_k| = 0 adds nothing to the product: both
// scales are 0, so it adds no rounding either (an empty operand gives an
// exactly zero sketch).  Rows whose norm is not finite keep the scales 1,
// so infinities and NaNs reach the sketch as without balancing.
//
// One wave per inner index k; the output is (scale of a_k, scale of b_k)
// in the element type, read once per k by the compress kernels.
#include <climits>

#include "device_common.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

constexpr int ROWS = 4;  // waves (inner indices) per workgroup

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// A lane's sum of squares as s 4^e.  float values: s in double, e = 0 (no
// square of a float overflows or underflows a double).  double values: each
// batch of U values is scaled by the power of two of its largest magnitude
// before squaring, and the batch sums are combined at the larger exponent;
// all scalings are exact powers of two.  bad: a value is infinite or NaN.
struct SumSq {
  double s = 0.0;
  int e = 0;
  bool bad = false;
};

template <int U, typename T>
__device__ __forceinline__ void add_squares(SumSq& a, const T (&v)[U]) {
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a.bad |= !(fabsf(v[u]) <= 3.4028234663852886e38f);
      a.s = fma(double(v[u]), double(v[u]), a.s);
    }
  } else {
    double mx = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a.bad |= !(fabs(v[u]) <= 1.7976931348623157e308);
      mx = fmax(mx, fabs(v[u]));
    }
    if (a.bad || mx == 0.0) return;
    const int ec = ilogb(mx);
    double cs = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double x = ldexp(double(v[u]), -ec);
      cs = fma(x, x, cs);
    }
    if (a.s == 0.0) {
      a.s = cs;
      a.e = ec;
    } else if (ec > a.e) {
      a.s = ldexp(a.s, 2 * (a.e - ec)) + cs;
      a.e = ec;
    } else {
      a.s += ldexp(cs, 2 * (ec - a.e));
    }
  }
}

// log2 of the Euclidean norm from the lanes' sums: returns 0 with *l2 set
// for a finite nonzero norm, 1 for a zero row, 2 for a non-finite one.
__device__ int wave_log2_norm(const SumSq& a, double* l2) {
  const bool bad = __any(a.bad);
  if (bad) return 2;
  const int e = wave_max_int(a.s > 0.0 ? a.e : INT_MIN);
  if (e == INT_MIN) return 1;
  const double s = wave_sum(a.s > 0.0 ? ldexp(a.s, 2 * (a.e - e)) : 0.0);
  if (!(s <= 1.7976931348623157e308)) return 2;  // float rows whose squares overflow the sum
  *l2 = double(e) + 0.5 * log2(s);
  return 0;
}

template <typename T>
__device__ __forceinline__ void pair_scale_body(const PairScaleArgs& a, int blk) {
  constexpr int U = 8;  // values of each row per lane with their loads in flight together
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6);
  const int lane = threadIdx.x & 63;
  const int k = blk * ROWS + w;
  if (k >= a.n) return;
  const T* ar = static_cast<const T*>(a.AT) + int64_t(k) * a.ld_at;
  const T* br = static_cast<const T*>(a.B) + int64_t(k) * a.ld_b;
  SumSq qa, qb;
  // rows of A and B in one pass, U values of each per lane and batch; the
  // index is clamped into the row, so every load is in bounds
  for (int i0 = 0; i0 < max(a.m, a.p); i0 += 64 * U) {
    T va[U], vb[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + 64 * u + lane;
      va[u] = a.m > 0 ? ar[min(i, a.m - 1)] : T(0);
      vb[u] = a.p > 0 ? br[min(i, a.p - 1)] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + 64 * u + lane;
      va[u] = i < a.m ? va[u] : T(0);
      vb[u] = i < a.p ? vb[u] : T(0);
    }
    add_squares<U>(qa, va);
    add_squares<U>(qb, vb);
  }
  double la = 0.0, lb = 0.0;
  const int sa = wave_log2_norm(qa, &la);
  const int sb = wave_log2_norm(qb, &lb);
  T ka = T(1), kb = T(1);
  if (sa == 1 || sb == 1) {
    ka = kb = T(0);
  } else if (sa == 0 && sb == 0) {
    constexpr double LIM = sizeof(T) == 4 ? 126.0 : 1000.0;  // 2^+-LIM stay normal
    const int e = int(fmin(LIM, fmax(-LIM, rint(0.5 * (lb - la)))));
    ka = T(ldexp(1.0, e));
    kb = T(ldexp(1.0, -e));
  }
  if (lane == 0) static_cast<cpx<T>*>(a.out)[k] = cpx<T>{ka, kb};
}

template <typename T>
__global__ void __launch_bounds__(64 * ROWS) k_pair_scale(PairScaleArgs a) {
  pair_scale_body<T>(a, blockIdx.x);
}

__global__ void __launch_bounds__(64 * ROWS) k_pair_scale_batch(const PairScaleArgs* __restrict__ jobs,
                                                                const int2* __restrict__ blocks) {
  const int2 e = blocks[blockIdx.x];
  pair_scale_body<float>(jobs[e.x], e.y);  // batches are FP32
}

}  // namespace

int pair_scale_blocks(const PairScaleArgs& a) { return (a.n + ROWS - 1) / ROWS; }

hipError_t launch_pair_scale(const PairScaleArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  const dim3 grid(unsigned(pair_scale_blocks(a)));
  if (a.f64)
    k_pair_scale<double><<<grid, dim3(64 * ROWS), 0, s>>>(a);
  else
    k_pair_scale<float><<<grid, dim3(64 * ROWS), 0, s>>>(a);
  return hipGetLastError();
}

hipError_t launch_pair_scale_batch(const PairScaleArgs* jobs, const int2* blocks, int nblocks,
                                   hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  k_pair_scale_batch<<<dim3(unsigned(nblocks)), dim3(64 * ROWS), 0, s>>>(jobs, blocks);
  return hipGetLastError();
}

}  // namespace cm
