// Twiddle tables of the FFT kernels, built on the device.
//
//   W_b^j = exp(-2 pi i j / b), j < b, in the element type: the angle is
//   reduced to the first octant with exact integer arithmetic (j = q b/8 + r),
//   cos and sin of 2 pi r / b or of its complement 2 pi (b/8 - r) / b are
//   taken in FP64 and placed by the octant's symmetry, then rounded once.
//   t2[q * 64 + L] = W_2048^(L * bitrev5(q)), the per-lane twiddles of the
//   2048-point register DFT (fft2048_reg.hpp), FP64 then rounded to FP32.
//
// The cached tables of the library and the tables the no-allocation entry
// writes into its scratch come from these kernels, so both give the same
// bits.
#include "device_common.hpp"
#include "kernels.hpp"

namespace cm {

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

__device__ __forceinline__ void cos_sin(int64_t num, int64_t b, double& c, double& s) {
  const double ang = kTwoPi * double(num) / double(b);
  c = cos(ang);
  s = sin(ang);
}

template <typename T>
__global__ void __launch_bounds__(256) k_twiddles(int lb, cpx<T>* __restrict__ out) {
  const int64_t b = int64_t(1) << lb;
  const int64_t j = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (j >= b) return;
  double c, s;
  if (b < 8) {
    cos_sin(j, b, c, s);
  } else {
    const int64_t q = (j * 8) >> lb;  // octant 0..7
    const int64_t r = j - q * (b >> 3);
    double c0, s0, c1, s1;
    cos_sin(r, b, c0, s0);            // angle in [0, pi/4)
    cos_sin((b >> 3) - r, b, c1, s1); // its complement to pi/4
    switch (q) {                      // cos, sin of q pi/4 + angle
      case 0: c = c0; s = s0; break;
      case 1: c = s1; s = c1; break;
      case 2: c = -s0; s = c0; break;
      case 3: c = -c1; s = s1; break;
      case 4: c = -c0; s = -s0; break;
      case 5: c = -s1; s = -c1; break;
      case 6: c = s0; s = -c0; break;
      default: c = c1; s = -s1; break;
    }
  }
  out[j] = cpx<T>{T(c), T(-s)};  // e^{-i angle}
}

__global__ void __launch_bounds__(256) k_t2_table(float2* __restrict__ out) {
  const int e = blockIdx.x * 256 + threadIdx.x;  // q * 64 + L
  if (e >= 2048) return;
  const int q = e >> 6, L = e & 63;
  const int br = int(__brev(uint32_t(q)) >> 27);  // bitrev5(q)
  const int j = (L * br) & 2047;
  double c, s;
  cos_sin(j, 2048, c, s);
  out[e] = make_float2(float(c), float(-s));
}

}  // namespace

size_t twiddle_bytes(int lb, int f64) { return (size_t(1) << lb) * 2 * (f64 ? 8 : 4); }

hipError_t launch_twiddles(int lb, int f64, void* out, hipStream_t s) {
  const int64_t b = int64_t(1) << lb;
  const dim3 grid(unsigned((b + 255) / 256));
  if (f64)
    k_twiddles<double><<<grid, dim3(256), 0, s>>>(lb, static_cast<cpx<double>*>(out));
  else
    k_twiddles<float><<<grid, dim3(256), 0, s>>>(lb, static_cast<cpx<float>*>(out));
  return hipGetLastError();
}

hipError_t launch_t2_table(void* out, hipStream_t s) {
  k_t2_table<<<dim3(8), dim3(256), 0, s>>>(static_cast<float2*>(out));
  return hipGetLastError();
}

}  // namespace cm
