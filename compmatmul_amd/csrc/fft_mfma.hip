// Matrix-core modes of the FFT compress (precision study paths): BF16 or
// FP16 operands with hi/lo split correction.
//
// The forward transforms of the pair-packed sketches z_k = S a_k + i S' b_k
// (length b = 16^S) run as S radix-16 Stockham stages whose 16-point DFTs
// are matrix products on the matrix cores: 32 DFT-16s at a time are
//   [Re y; Im y] (32 x 32) = M (32 x 32) [Re x; Im x] (32 x 32),
//   M = [[C, S], [-S, C]],  C[a][r] = cos(2 pi a r / 16), S[a][r] = sin(...),
// i.e. two K = 16 steps of v_mfma_f32_32x32x16_bf16 (or _f16).  Operands are
// split into hi + lo parts of the 16-bit type and each product is
// hi*hi + hi*lo + lo*hi with FP32 accumulation; the lo*lo term is dropped.
//  * BF16 (8-bit significand): x = hi + lo to within 2^-17 |x|; the dropped
//    term is ~2^-16 relative.
//  * FP16 (11-bit significand, largest finite 65504): each DFT-16 input
//    column is first scaled by a power of two 2^s so that its largest
//    magnitude lies in [2^13, 2^14) (exact, and far from overflow and from
//    the subnormal range), split to within 2^-23 of that largest magnitude,
//    and the outputs are scaled back by 2^-s; the dropped term is ~2^-22.
// Twiddles between stages are applied in FP32 from the W_b table.
//
// Pipeline per repetition t and chunk of rows: zero the rows, scatter the
// sketches (FP32 atomics in HBM), S stages ping-ponging between two row
// buffers, then Q = sum_k z_k^2 written into the sliced spectrum layout
// [t][F mod b1][F / b1] (b1 = b / 2048) that the reduce and combine kernels
// of the fused path finish.
//
// Stockham stage s (Ns = 16^s): for j < b/16 with kk = j mod Ns,
//   v_r = x[j + r b/16] * W_b^(r kk b / (16 Ns)),  v = DFT16(v),
//   y[(j / Ns) 16 Ns + kk + a Ns] = v_a.
// Operand layouts of the 32x32x16 MFMA (both types; lane l, i = l mod 32, h = l/32):
//   A: A[i][8h + q], B: B[8h + q][i], q < 8; D register r: D[8(r/4) + 4h + r mod 4][i].
#include <algorithm>

#include "device_common.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// cos and sin of 2 pi t / 16, rounded from double
constexpr float C16[16] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                           0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                           -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
                           0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
constexpr float S16[16] = {0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
                           1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
                           0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
                           -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f};

template <typename H>
__device__ __forceinline__ void split(float x, H& hi, H& lo) {
  hi = (H)x;
  lo = (H)(x - (float)hi);
}

template <typename H>
struct Vec8;
template <>
struct Vec8<__bf16> {
  typedef bf16x8 type;
};
template <>
struct Vec8<_Float16> {
  typedef f16x8 type;
};

template <typename H>
__device__ __forceinline__ f32x16 mfma32(typename Vec8<H>::type a, typename Vec8<H>::type b, f32x16 c) {
  if constexpr (__is_same(H, __bf16))
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// FP16 column scale: the power of two 2^s with max * 2^s in [2^13, 2^14)
// (1 for a zero or non-finite column), as its exponent s
__device__ __forceinline__ int f16_scale_exp(float maxabs) {
  if (!(maxabs > 0.f) || !__builtin_isfinite(maxabs)) return 0;
  return 13 - ilogbf(maxabs);
}

// element (i, c) of the real 32 x 32 DFT-16 matrix
__device__ __forceinline__ float mat(int i, int c) {
  const int a = i & 15, r = c & 15, t = (a * r) & 15;
  if (i < 16) return c < 16 ? C16[t] : S16[t];
  return c < 16 ? -S16[t] : C16[t];
}

__global__ void __launch_bounds__(256) k_zero_rows(float4* __restrict__ z, int64_t n4) {
  for (int64_t e = int64_t(blockIdx.x) * 256 + threadIdx.x; e < n4; e += int64_t(gridDim.x) * 256)
    z[e] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// pair-packed sketch of row k0 + blockIdx.x into row blockIdx.x of Z, the
// rows times their power-of-two balance scales (balance.hip)
__global__ void __launch_bounds__(256) k_sketch_rows(const float* __restrict__ AT, int64_t ld_at,
                                                     const float* __restrict__ B, int64_t ld_b,
                                                     const uint32_t* __restrict__ hA,
                                                     const uint32_t* __restrict__ hB, int m, int p,
                                                     int k0, int lb, const float2* __restrict__ scale,
                                                     float2* __restrict__ Z) {
  const int r = blockIdx.x;
  float2* z = Z + (int64_t(r) << lb);
  const float* ar = AT + int64_t(k0 + r) * ld_at;
  const float* br = B + int64_t(k0 + r) * ld_b;
  const float2 bal = scale[k0 + r];
  for (int i = threadIdx.x; i < m; i += 256) {
    const uint32_t e = hA[i];
    const float v = ar[i] * bal.x;
    atomicAdd(&z[hbucket(e)].x, hneg(e) ? -v : v);
  }
  for (int j = threadIdx.x; j < p; j += 256) {
    const uint32_t e = hB[j];
    const float v = br[j] * bal.y;
    atomicAdd(&z[hbucket(e)].y, hneg(e) ? -v : v);
  }
}

// one radix-16 Stockham stage over `rows` rows of length 2^lb; each wave
// transforms 32 consecutive j of one row
template <typename H>
__global__ void __launch_bounds__(256) k_fft16_stage(const float2* __restrict__ X,
                                                     float2* __restrict__ Y, int lb, int ns_log,
                                                     const float2* __restrict__ tw, int rows) {
  typedef typename Vec8<H>::type v8;
  constexpr bool F16 = !__is_same(H, __bf16);
  const int lane = threadIdx.x & 63;
  const int i = lane & 31, h = lane >> 5;
  const int64_t N = int64_t(1) << lb;
  const int waves_per_row = int(N >> 9);  // (N / 16) / 32
  const int gw = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (gw >= rows * waves_per_row) return;  // whole waves: the lane exchange below stays within one
  const int row = gw / waves_per_row;
  const int j = (gw - row * waves_per_row) * 32 + i;
  const int Ns = 1 << ns_log;
  const int kk = j & (Ns - 1);
  const float2* x = X + (int64_t(row) << lb);
  float2* y = Y + (int64_t(row) << lb);
  const int64_t quarter = N >> 4;  // b / 16

  // this lane's DFT-matrix operands (A: row i, columns 8h..8h+7 and 16+8h..)
  v8 a0h, a0l, a1h, a1l;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    H hi, lo;
    split(mat(i, 8 * h + q), hi, lo);
    a0h[q] = hi;
    a0l[q] = lo;
    split(mat(i, 16 + 8 * h + q), hi, lo);
    a1h[q] = hi;
    a1l[q] = lo;
  }
  // inputs r = 8h .. 8h+7 of column j, twiddled in FP32
  float2 v[8];
  const int tw_shift = lb - 4 - ns_log;  // b / (16 Ns) = 2^tw_shift
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int r = 8 * h + q;
    v[q] = x[j + r * quarter];
    if (ns_log > 0) {
      const float2 w = tw[int64_t(r * kk) << tw_shift];
      v[q] = make_float2(v[q].x * w.x - v[q].y * w.y, v[q].x * w.y + v[q].y * w.x);
    }
  }
  // FP16: the column (lanes i and i + 32) scaled by 2^s before the split
  int s = 0;
  if constexpr (F16) {
    float mx = 0.f;
#pragma unroll
    for (int q = 0; q < 8; ++q) mx = fmaxf(mx, fmaxf(fabsf(v[q].x), fabsf(v[q].y)));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    s = f16_scale_exp(mx);
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = make_float2(ldexpf(v[q].x, s), ldexpf(v[q].y, s));
  }
  v8 reh, rel, imh, iml;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    H hi, lo;
    split(v[q].x, hi, lo);
    reh[q] = hi;
    rel[q] = lo;
    split(v[q].y, hi, lo);
    imh[q] = hi;
    iml[q] = lo;
  }
  f32x16 acc = {};
  acc = mfma32<H>(a0l, reh, acc);
  acc = mfma32<H>(a0h, rel, acc);
  acc = mfma32<H>(a1l, imh, acc);
  acc = mfma32<H>(a1h, iml, acc);
  acc = mfma32<H>(a0h, reh, acc);
  acc = mfma32<H>(a1h, imh, acc);
  // register g holds output row 8(g/4) + 4h + g%4 of column i: rows < 16
  // real parts of frequency a, rows >= 16 imaginary parts of a - 16
  const int64_t base = int64_t(j >> ns_log) * (int64_t(Ns) << 4) + kk;
#pragma unroll
  for (int g = 0; g < 8; ++g) {
    const int a = 8 * (g >> 2) + 4 * h + (g & 3);  // g < 8: real rows
    float2 o = make_float2(acc[g], acc[g + 8]);
    if constexpr (F16) o = make_float2(ldexpf(o.x, -s), ldexpf(o.y, -s));
    y[base + int64_t(a) * Ns] = o;
  }
}

// Q[F] = sum over the rows of z[F]^2, into the sliced layout of repetition t
__global__ void __launch_bounds__(256) k_mac_sliced(const float2* __restrict__ Z, int rows, int lb,
                                                    float2* __restrict__ out) {
  const int64_t N = int64_t(1) << lb;
  const int64_t F = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (F >= N) return;
  float sx = 0.f, sy = 0.f;
  for (int r = 0; r < rows; ++r) {
    const float2 z = Z[(int64_t(r) << lb) + F];
    sx += z.x * z.x - z.y * z.y;
    sy = fmaf(2.f * z.x, z.y, sy);
  }
  const int lbs = lb - kRegLb2;
  const int64_t sl = F & ((int64_t(1) << lbs) - 1), f = F >> lbs;
  out[(sl << kRegLb2) + f] = make_float2(sx, sy);
}

}  // namespace

int fft_bf16_supported(int lb) { return lb >= 12 && lb <= 28 && lb % 4 == 0; }

hipError_t launch_fft_bf16_compress(const Bf16FftArgs& a, hipStream_t s) {
  const int64_t N = int64_t(1) << a.lb;
  const int stages = a.lb / 4;
  for (int t = 0; t < a.d; ++t) {
    const uint32_t* hA = a.hA + int64_t(t) * a.m;
    const uint32_t* hB = a.hB + int64_t(t) * a.p;
    for (int c = 0; c < a.nchunks; ++c) {
      const int k0 = c * a.rows;
      const int rows = min(a.rows, a.n - k0);
      float2* bufs[2] = {a.Z0, a.Z1};
      const int64_t n4 = (int64_t(rows) << a.lb) / 2;  // float4 per 2 complex
      k_zero_rows<<<dim3(unsigned(std::min<int64_t>(4096, (n4 + 255) / 256))), dim3(256), 0, s>>>(
          reinterpret_cast<float4*>(bufs[0]), n4);
      k_sketch_rows<<<dim3(unsigned(rows)), dim3(256), 0, s>>>(a.AT, a.ld_at, a.B, a.ld_b, hA, hB, a.m,
                                                              a.p, k0, a.lb, a.scale, bufs[0]);
      const int waves = int(rows * (N >> 9));
      for (int st = 0; st < stages; ++st)
        if (a.fp16)
          k_fft16_stage<_Float16><<<dim3(unsigned((waves + 3) / 4)), dim3(256), 0, s>>>(
              bufs[st & 1], bufs[(st + 1) & 1], a.lb, 4 * st, static_cast<const float2*>(a.tw), rows);
        else
          k_fft16_stage<__bf16><<<dim3(unsigned((waves + 3) / 4)), dim3(256), 0, s>>>(
              bufs[st & 1], bufs[(st + 1) & 1], a.lb, 4 * st, static_cast<const float2*>(a.tw), rows);
      float2* out = reinterpret_cast<float2*>(a.partial) + (int64_t(c) * a.d + t) * N;
      k_mac_sliced<<<dim3(unsigned((N + 255) / 256)), dim3(256), 0, s>>>(bufs[stages & 1], rows, a.lb,
                                                                        out);
    }
  }
  return hipGetLastError();
}

}  // namespace cm
