// Unfused "plain" GPU pipeline and the FP64 GEMM.
//
// The plain path runs the reference's stages as separate launches per
// repetition t (matmul_fft.hpp:70-107, matmul_fwht.hpp:64-95): count-sketch
// every row k into HBM, transform every sketch in place, then reduce the
// pointwise products over k.  It materialises n x b sketches and spectra in
// HBM and is the per-product baseline the fused path is measured against.
// It uses the same pair packing (z = S a + i S' b, accumulate z^2) and the
// same reduce/combine kernels as the fused path, so the two differ only in
// fusion.  It handles transforms that fit 64 KB of LDS: b <= 2^14 (FWHT, float),
// 2^13 (FWHT double, FFT float), 2^12 (FFT double).
//
// cm::launch_gemm_f64 is the exact FP64 product behind compmatmul.gemm
// (array.hpp:81-95: C = AT^T B, all row-major).
#include "device_common.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

constexpr int NT = 256;

template <typename T, bool FWHT>
__global__ void __launch_bounds__(NT) k_plain_sketch(PlainArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = 1 << a.lb;
  const int k = blockIdx.x;
  const int tid = threadIdx.x;
  const T* __restrict__ ar = static_cast<const T*>(a.AT) + int64_t(k) * a.ld_at;
  const T* __restrict__ br = static_cast<const T*>(a.B) + int64_t(k) * a.ld_b;
  const uint32_t* __restrict__ hA = a.hA + int64_t(a.t) * a.m;
  const uint32_t* __restrict__ hB = a.hB + int64_t(a.t) * a.p;
  if constexpr (FWHT) {
    // blockIdx.y selects the A side (0) or the B side (1)
    T* s = reinterpret_cast<T*>(smem_raw);
    const bool side_b = blockIdx.y != 0;
    const int keys = side_b ? a.p : a.m;
    const T* __restrict__ row = side_b ? br : ar;
    const uint32_t* __restrict__ hh = side_b ? hB : hA;
    for (int e = tid; e < b; e += NT) s[e] = T(0);
    __syncthreads();
    for (int i = tid; i < keys; i += NT) atomicAdd(&s[hbucket(hh[i])], apply_sign(row[i], hneg(hh[i])));
    __syncthreads();
    T* S = static_cast<T*>(side_b ? a.SB : a.SA) + int64_t(k) * b;
    for (int e = tid; e < b; e += NT) S[e] = s[e];
  } else {
    cpx<T>* z = reinterpret_cast<cpx<T>*>(smem_raw);
    const cpx<T> bal = static_cast<const cpx<T>*>(a.scale)[k];  // balance.hip
    for (int e = tid; e < b; e += NT) z[e] = {T(0), T(0)};
    __syncthreads();
    for (int i = tid; i < a.m; i += NT)
      atomicAdd(&z[hbucket(hA[i])].x, apply_sign(ar[i] * bal.x, hneg(hA[i])));
    for (int j = tid; j < a.p; j += NT)
      atomicAdd(&z[hbucket(hB[j])].y, apply_sign(br[j] * bal.y, hneg(hB[j])));
    __syncthreads();
    cpx<T>* Z = static_cast<cpx<T>*>(a.SA) + int64_t(k) * b;
    for (int e = tid; e < b; e += NT) Z[e] = z[e];
  }
}

template <typename T, bool FWHT>
__global__ void __launch_bounds__(NT) k_plain_transform(PlainArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = 1 << a.lb;
  const int k = blockIdx.x;
  const int tid = threadIdx.x;
  if constexpr (FWHT) {
    T* s = reinterpret_cast<T*>(smem_raw);
    T* V = static_cast<T*>(blockIdx.y ? a.SB : a.SA) + int64_t(k) * b;
    for (int e = tid; e < b; e += NT) s[e] = V[e];
    __syncthreads();
    for (int len = 1; len < b; len <<= 1) {
      for (int q = tid; q < (b >> 1); q += NT) {
        const int j = ((q & ~(len - 1)) << 1) | (q & (len - 1));
        const T x0 = s[j], x1 = s[j + len];
        s[j] = x0 + x1;
        s[j + len] = x0 - x1;
      }
      __syncthreads();
    }
    for (int e = tid; e < b; e += NT) V[e] = s[e];
  } else {
    cpx<T>* z = reinterpret_cast<cpx<T>*>(smem_raw);
    const cpx<T>* __restrict__ tw = static_cast<const cpx<T>*>(a.tw);
    cpx<T>* Z = static_cast<cpx<T>*>(a.SA) + int64_t(k) * b;
    for (int e = tid; e < b; e += NT) z[e] = Z[e];
    __syncthreads();
    for (int len = b >> 1; len >= 1; len >>= 1) {
      const int lstride = a.lb - __ffs(len);
      for (int q = tid; q < (b >> 1); q += NT) {
        const int pos = q & (len - 1);
        const int j = ((q & ~(len - 1)) << 1) | pos;
        const cpx<T> x0 = z[j], x1 = z[j + len];
        z[j] = cadd(x0, x1);
        z[j + len] = cmul(csub(x0, x1), tw[int64_t(pos) << lstride]);
      }
      __syncthreads();
    }
    for (int e = tid; e < b; e += NT) Z[e] = z[e];  // bit-reversed frequency order
  }
}

// acc[f] = sum_k A_k[f] B_k[f] (FWHT) or sum_k Z_k[f]^2 (FFT, stored at the
// natural frequency index); acc uses the [t][slice=0][b] partial layout.
template <typename T, bool FWHT>
__global__ void __launch_bounds__(NT) k_plain_mac(PlainArgs a) {
  const int b = 1 << a.lb;
  const int f = blockIdx.x * NT + threadIdx.x;
  if (f >= b) return;
  if constexpr (FWHT) {
    const T* SA = static_cast<const T*>(a.SA);
    const T* SB = static_cast<const T*>(a.SB);
    T s = T(0);
    for (int k = 0; k < a.n; ++k) s += SA[int64_t(k) * b + f] * SB[int64_t(k) * b + f];
    static_cast<T*>(a.acc)[f] = s;
  } else {
    const cpx<T>* Z = static_cast<const cpx<T>*>(a.SA);
    cpx<T> s = {T(0), T(0)};
    for (int k = 0; k < a.n; ++k) s = cadd(s, csqr(Z[int64_t(k) * b + f]));
    static_cast<cpx<T>*>(a.acc)[bitrev(f, a.lb)] = s;
  }
}

template <typename T, bool FWHT>
hipError_t plain_sketch_t(const PlainArgs& a, hipStream_t s) {
  const size_t lds = (FWHT ? 1 : 2) * sizeof(T) * (size_t(1) << a.lb);
  k_plain_sketch<T, FWHT><<<dim3(unsigned(a.n), FWHT ? 2 : 1), dim3(NT), lds, s>>>(a);
  return hipGetLastError();
}

template <typename T, bool FWHT>
hipError_t plain_transform_t(const PlainArgs& a, hipStream_t s) {
  const size_t lds = (FWHT ? 1 : 2) * sizeof(T) * (size_t(1) << a.lb);
  k_plain_transform<T, FWHT><<<dim3(unsigned(a.n), FWHT ? 2 : 1), dim3(NT), lds, s>>>(a);
  return hipGetLastError();
}

template <typename T, bool FWHT>
hipError_t plain_mac_t(const PlainArgs& a, hipStream_t s) {
  const int b = 1 << a.lb;
  k_plain_mac<T, FWHT><<<dim3(unsigned((b + NT - 1) / NT)), dim3(NT), 0, s>>>(a);
  return hipGetLastError();
}

// C[i][j] = sum_k AT[k][i] B[k][j], 16x16 output tiles staged through LDS
__global__ void __launch_bounds__(256) k_gemm_f64(int m, int n, int p,
                                                  const double* __restrict__ AT,
                                                  const double* __restrict__ B,
                                                  double* __restrict__ C) {
  __shared__ double sa[16][17], sb[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int i = blockIdx.y * 16 + ty, j = blockIdx.x * 16 + tx;
  double acc = 0.0;
  for (int k0 = 0; k0 < n; k0 += 16) {
    const int ka = k0 + ty;
    sa[ty][tx] = (ka < n && blockIdx.y * 16 + tx < m) ? AT[int64_t(ka) * m + blockIdx.y * 16 + tx] : 0.0;
    sb[ty][tx] = (ka < n && j < p) ? B[int64_t(ka) * p + j] : 0.0;
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) acc = fma(sa[kk][ty], sb[kk][tx], acc);
    __syncthreads();
  }
  if (i < m && j < p) C[int64_t(i) * p + j] = acc;
}

}  // namespace

hipError_t launch_plain_sketch(const PlainArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.f64) return a.fwht ? plain_sketch_t<double, true>(a, s) : plain_sketch_t<double, false>(a, s);
  return a.fwht ? plain_sketch_t<float, true>(a, s) : plain_sketch_t<float, false>(a, s);
}

hipError_t launch_plain_transform(const PlainArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  if (a.f64) return a.fwht ? plain_transform_t<double, true>(a, s) : plain_transform_t<double, false>(a, s);
  return a.fwht ? plain_transform_t<float, true>(a, s) : plain_transform_t<float, false>(a, s);
}

hipError_t launch_plain_mac(const PlainArgs& a, hipStream_t s) {
  if (a.f64) return a.fwht ? plain_mac_t<double, true>(a, s) : plain_mac_t<double, false>(a, s);
  return a.fwht ? plain_mac_t<float, true>(a, s) : plain_mac_t<float, false>(a, s);
}

hipError_t launch_gemm_f64(int m, int n, int p, const double* AT, const double* B, double* C,
                           hipStream_t s) {
  if (m == 0 || p == 0) return hipSuccess;
  const dim3 grid(unsigned((p + 15) / 16), unsigned((m + 15) / 16));
  k_gemm_f64<<<grid, dim3(256), 0, s>>>(m, n, p, AT, B, C);
  return hipGetLastError();
}

}  // namespace cm
