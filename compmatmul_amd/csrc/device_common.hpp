// Device-side helpers shared by the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cm {

template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> csqr(cpx<T> a) {
  return {a.x * a.x - a.y * a.y, T(2) * a.x * a.y};
}

// Packed hash entry: bucket in bits 0..30, bit 31 set when the sign is -1.
__device__ __forceinline__ uint32_t hbucket(uint32_t e) { return e & 0x7fffffffu; }
__device__ __forceinline__ uint32_t hneg(uint32_t e) { return e >> 31; }

template <typename T>
__device__ __forceinline__ T apply_sign(T v, uint32_t neg) { return neg ? -v : v; }

// reverse the low `bits` bits of x (bits >= 1)
__device__ __forceinline__ uint32_t bitrev(uint32_t x, int bits) {
  return __brev(x) >> (32 - bits);
}

}  // namespace cm
