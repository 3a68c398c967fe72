// 2048-point complex DFT of a wave's LDS-resident sketch, held in registers
// (shared by the FFT register kernels: compress_fft_reg.hip and
// compress_fft_grp.hip).
//
// Element n = L + 64 r (lane L, register r < 32) as read from LDS.  Four-step
// DFT with n = n1 + 64 n2, k = k2 + 32 k1:
//   1. per lane, DFT-32 over the registers (DIF: one radix-2^2 step, three
//      radix-2 stages, compile-time twiddles; register q then holds
//      k2 = bitrev5(q));
//   2. twiddle by W_2048^(n1 k2) (per-lane table, L1-resident);
//   3. transpose through LDS (rows of 66 complex: 16-byte aligned, read
//      as conflict-free b128 pairs from one base address plus
//      immediates): lane L' gets k2 = L' & 31 and the 32 values
//      n1 = r' + 32 (L' >> 5);
//   4. DFT-64 over n1: one radix-2 DIF stage across lanes L' and L'+32 with
//      v_permlane32_swap, a second swap that regroups the halves, then a
//      DFT-32 per lane (lanes < 32 give even k1, lanes >= 32 odd).
#pragma once
#include "device_common.hpp"

namespace cm {
namespace f2k {

constexpr int R = 32;     // complex values per lane
// transpose rows of 64 complex padded to 66 (32 x 66 = 2112 slots, the
// partition size): 16-byte aligned, so the transposed read is b128 pairs,
// each 8-lane pass on 32 distinct banks; the row writes stay conflict-free
constexpr int TROW = 66;

// exp(-2 pi i j / N), rounded to float
constexpr float W32R[32] = {1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f, 8.314696123e-01f, 7.071067812e-01f, 5.555702330e-01f, 3.826834324e-01f, 1.950903220e-01f, 0.0f, -1.950903220e-01f, -3.826834324e-01f, -5.555702330e-01f, -7.071067812e-01f, -8.314696123e-01f, -9.238795325e-01f, -9.807852804e-01f, -1.000000000e+00f, -9.807852804e-01f, -9.238795325e-01f, -8.314696123e-01f, -7.071067812e-01f, -5.555702330e-01f, -3.826834324e-01f, -1.950903220e-01f, 0.0f, 1.950903220e-01f, 3.826834324e-01f, 5.555702330e-01f, 7.071067812e-01f, 8.314696123e-01f, 9.238795325e-01f, 9.807852804e-01f};
constexpr float W32I[32] = {0.0f, -1.950903220e-01f, -3.826834324e-01f, -5.555702330e-01f, -7.071067812e-01f, -8.314696123e-01f, -9.238795325e-01f, -9.807852804e-01f, -1.000000000e+00f, -9.807852804e-01f, -9.238795325e-01f, -8.314696123e-01f, -7.071067812e-01f, -5.555702330e-01f, -3.826834324e-01f, -1.950903220e-01f, 0.0f, 1.950903220e-01f, 3.826834324e-01f, 5.555702330e-01f, 7.071067812e-01f, 8.314696123e-01f, 9.238795325e-01f, 9.807852804e-01f, 1.000000000e+00f, 9.807852804e-01f, 9.238795325e-01f, 8.314696123e-01f, 7.071067812e-01f, 5.555702330e-01f, 3.826834324e-01f, 1.950903220e-01f};
constexpr float W64R[64] = {1.000000000e+00f, 9.951847267e-01f, 9.807852804e-01f, 9.569403357e-01f, 9.238795325e-01f, 8.819212643e-01f, 8.314696123e-01f, 7.730104534e-01f, 7.071067812e-01f, 6.343932842e-01f, 5.555702330e-01f, 4.713967368e-01f, 3.826834324e-01f, 2.902846773e-01f, 1.950903220e-01f, 9.801714033e-02f, 0.0f, -9.801714033e-02f, -1.950903220e-01f, -2.902846773e-01f, -3.826834324e-01f, -4.713967368e-01f, -5.555702330e-01f, -6.343932842e-01f, -7.071067812e-01f, -7.730104534e-01f, -8.314696123e-01f, -8.819212643e-01f, -9.238795325e-01f, -9.569403357e-01f, -9.807852804e-01f, -9.951847267e-01f, -1.000000000e+00f, -9.951847267e-01f, -9.807852804e-01f, -9.569403357e-01f, -9.238795325e-01f, -8.819212643e-01f, -8.314696123e-01f, -7.730104534e-01f, -7.071067812e-01f, -6.343932842e-01f, -5.555702330e-01f, -4.713967368e-01f, -3.826834324e-01f, -2.902846773e-01f, -1.950903220e-01f, -9.801714033e-02f, 0.0f, 9.801714033e-02f, 1.950903220e-01f, 2.902846773e-01f, 3.826834324e-01f, 4.713967368e-01f, 5.555702330e-01f, 6.343932842e-01f, 7.071067812e-01f, 7.730104534e-01f, 8.314696123e-01f, 8.819212643e-01f, 9.238795325e-01f, 9.569403357e-01f, 9.807852804e-01f, 9.951847267e-01f};
constexpr float W64I[64] = {0.0f, -9.801714033e-02f, -1.950903220e-01f, -2.902846773e-01f, -3.826834324e-01f, -4.713967368e-01f, -5.555702330e-01f, -6.343932842e-01f, -7.071067812e-01f, -7.730104534e-01f, -8.314696123e-01f, -8.819212643e-01f, -9.238795325e-01f, -9.569403357e-01f, -9.807852804e-01f, -9.951847267e-01f, -1.000000000e+00f, -9.951847267e-01f, -9.807852804e-01f, -9.569403357e-01f, -9.238795325e-01f, -8.819212643e-01f, -8.314696123e-01f, -7.730104534e-01f, -7.071067812e-01f, -6.343932842e-01f, -5.555702330e-01f, -4.713967368e-01f, -3.826834324e-01f, -2.902846773e-01f, -1.950903220e-01f, -9.801714033e-02f, 0.0f, 9.801714033e-02f, 1.950903220e-01f, 2.902846773e-01f, 3.826834324e-01f, 4.713967368e-01f, 5.555702330e-01f, 6.343932842e-01f, 7.071067812e-01f, 7.730104534e-01f, 8.314696123e-01f, 8.819212643e-01f, 9.238795325e-01f, 9.569403357e-01f, 9.807852804e-01f, 9.951847267e-01f, 1.000000000e+00f, 9.951847267e-01f, 9.807852804e-01f, 9.569403357e-01f, 9.238795325e-01f, 8.819212643e-01f, 8.314696123e-01f, 7.730104534e-01f, 7.071067812e-01f, 6.343932842e-01f, 5.555702330e-01f, 4.713967368e-01f, 3.826834324e-01f, 2.902846773e-01f, 1.950903220e-01f, 9.801714033e-02f};


// x W_32^e for an exponent known after unrolling: quarter turns are exact
// moves, the rest one complex product
__device__ __forceinline__ void twiddle32(float& xr, float& xi, int e) {
  e &= 31;
  if (e == 0) return;
  if (e == 8) {  // -i
    const float t = xr;
    xr = xi;
    xi = -t;
    return;
  }
  if (e == 16) {
    xr = -xr;
    xi = -xi;
    return;
  }
  if (e == 24) {  // +i
    const float t = xr;
    xr = -xi;
    xi = t;
    return;
  }
  const float wr = W32R[e], wi = W32I[e];
  const float nr = fmaf(xr, wr, -xi * wi);
  xi = fmaf(xr, wi, xi * wr);
  xr = nr;
}

// in-register DIF DFT-32: register q ends holding X[bitrev5(q)].  The
// first two radix-2 stages run as one radix-2^2 step: for a, b, c, d =
// x[j], x[j+8], x[j+16], x[j+24], p = a + c, q = b + d, r = a - c, t = b - d,
// the outputs are p + q, (p - q) W^2j, (r - i t) W^j and (r + i t) W^3j,
// 20 general twiddle products where two radix-2 stages take 26.
__device__ __forceinline__ void dft32(float (&re)[R], float (&im)[R]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const float pr = re[j] + re[j + 16], pi = im[j] + im[j + 16];
    const float rr = re[j] - re[j + 16], ri = im[j] - im[j + 16];
    const float qr = re[j + 8] + re[j + 24], qi = im[j + 8] + im[j + 24];
    const float tr = re[j + 8] - re[j + 24], ti = im[j + 8] - im[j + 24];
    re[j] = pr + qr;
    im[j] = pi + qi;
    float ar = pr - qr, ai = pi - qi;
    twiddle32(ar, ai, 2 * j);
    re[j + 8] = ar;
    im[j + 8] = ai;
    float br = rr + ti, bi = ri - tr;  // r - i t
    twiddle32(br, bi, j);
    re[j + 16] = br;
    im[j + 16] = bi;
    float cr = rr - ti, ci = ri + tr;  // r + i t
    twiddle32(cr, ci, 3 * j);
    re[j + 24] = cr;
    im[j + 24] = ci;
  }
#pragma unroll
  for (int half = R / 8; half >= 1; half >>= 1) {
#pragma unroll
    for (int blk = 0; blk < R; blk += 2 * half) {
#pragma unroll
      for (int j = 0; j < half; ++j) {
        const int a = blk + j, b = a + half;
        const float ur = re[a] + re[b], ui = im[a] + im[b];
        const float vr = re[a] - re[b], vi = im[a] - im[b];
        re[a] = ur;
        im[a] = ui;
        const int tj = j * (R / (2 * half));  // W_{2 half}^j = W_32^tj
        if (tj == 0) {
          re[b] = vr;
          im[b] = vi;
        } else if (tj == R / 4) {  // -i
          re[b] = vi;
          im[b] = -vr;
        } else {
          const float wr = W32R[tj], wi = W32I[tj];
          re[b] = fmaf(vr, wr, -vi * wi);
          im[b] = fmaf(vr, wi, vi * wr);
        }
      }
    }
  }
}

__device__ __forceinline__ int brev5(int q) { return int(__brev(uint32_t(q)) >> 27); }
__host__ __device__ constexpr int cbrev5(int q) {
  return ((q & 1) << 4) | ((q & 2) << 2) | (q & 4) | ((q & 8) >> 2) | ((q & 16) >> 4);
}

// Ordering of this wave's LDS accesses between the phases: a workgroup
// barrier, or (when only this wave touches S) a compiler barrier, since the
// LDS operations of one wave complete in order.
template <bool WAVE_LOCAL>
__device__ __forceinline__ void order() {
  if constexpr (WAVE_LOCAL)
    asm volatile("" ::: "memory");
  else
    __syncthreads();
}

// DFT of the 2048 complex values S[0 .. 2048) of this wave; S (2112 slots)
// is left dirty.  On return lane L, register q holds frequency fq_of(L, q).
template <bool WAVE_LOCAL>
__device__ __forceinline__ void dft2048_lds(float2* S, const float2* __restrict__ t2, int lane,
                                            float (&re)[R], float (&im)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const float2 z = S[r * 64 + lane];
    re[r] = z.x;
    im[r] = z.y;
  }
  order<WAVE_LOCAL>();
  dft32(re, im);  // over n2; register q holds k2 = bitrev5(q)
  // twiddle W_2048^(n1 k2), k2 = j + 8 i: the table (row bitrev5(k2) holds
  // W_2048^(n1 k2)) gives A[j] = W^(n1 j) and B[i] = W^(n1 8 i), ten loads
  // per call instead of 32; the other twiddles are products A[j] B[i].  The
  // table address is laundered so the compiler reloads these (L1-resident)
  // instead of pinning them across the k loop.
  asm volatile("" : "+s"(t2));
  float2 A[8], B[4];
  A[0] = make_float2(1.f, 0.f);
  B[0] = A[0];
#pragma unroll
  for (int jj = 1; jj < 8; ++jj) A[jj] = (t2 + cbrev5(jj) * 64)[uint32_t(lane)];
#pragma unroll
  for (int ii = 1; ii < 4; ++ii) B[ii] = (t2 + cbrev5(8 * ii) * 64)[uint32_t(lane)];
#pragma unroll
  for (int q = 1; q < R; ++q) {  // q = 0 is k2 = 0: W = 1
    const int k2 = cbrev5(q), jj = k2 & 7, ii = k2 >> 3;
    float2 w;
    if (ii == 0)
      w = A[jj];
    else if (jj == 0)
      w = B[ii];
    else
      w = make_float2(fmaf(A[jj].x, B[ii].x, -A[jj].y * B[ii].y), fmaf(A[jj].x, B[ii].y, A[jj].y * B[ii].x));
    const float xr = re[q], xi = im[q];
    re[q] = fmaf(xr, w.x, -xi * w.y);
    im[q] = fmaf(xr, w.y, xi * w.x);
  }
  // transpose: (n1 = L, k2) -> row k2, column n1
#pragma unroll
  for (int q = 0; q < R; ++q) S[brev5(q) * TROW + lane] = make_float2(re[q], im[q]);
  order<WAVE_LOCAL>();
  {
    const float4* rbase = reinterpret_cast<const float4*>(S + (lane & 31) * TROW + 32 * (lane >> 5));
#pragma unroll
    for (int r = 0; r < R; r += 2) {
      const float4 z = rbase[r / 2];
      re[r] = z.x;
      im[r] = z.y;
      re[r + 1] = z.z;
      im[r + 1] = z.w;
    }
  }
  order<WAVE_LOCAL>();
  // radix-2 DIF stage over n1 bit 5 (lane bit 5): pairs (r', lane) and
  // (r', lane ^ 32); lanes < 32 finish r' = 2j, lanes >= 32 r' = 2j + 1
  // (the lane half is laundered so the per-lane twiddle selection is redone
  // here rather than held in 64 VGPRs across the k loop)
  int lh = lane;
  asm volatile("" : "+v"(lh));
  const bool hi = lh >= 32;
#pragma unroll
  for (int j = 0; j < R / 2; ++j) {
    const auto pr = __builtin_amdgcn_permlane32_swap(__float_as_uint(re[2 * j]),
                                                     __float_as_uint(re[2 * j + 1]), false, false);
    const auto pi = __builtin_amdgcn_permlane32_swap(__float_as_uint(im[2 * j]),
                                                     __float_as_uint(im[2 * j + 1]), false, false);
    const float ar_ = __uint_as_float(pr[0]), br_ = __uint_as_float(pr[1]);
    const float ai_ = __uint_as_float(pi[0]), bi_ = __uint_as_float(pi[1]);
    const float vr = ar_ - br_, vi = ai_ - bi_;
    const float wr = hi ? W64R[2 * j + 1] : W64R[2 * j];
    const float wi = hi ? W64I[2 * j + 1] : W64I[2 * j];
    re[2 * j] = ar_ + br_;  // even-k1 input at r' = 2j + hi
    im[2 * j] = ai_ + bi_;
    re[2 * j + 1] = fmaf(vr, wr, -vi * wi);  // odd-k1 input at r' = 2j + hi
    im[2 * j + 1] = fmaf(vr, wi, vi * wr);
  }
  // regroup: lanes < 32 collect the even-k1 inputs for all r', lanes >= 32
  // the odd-k1 inputs
#pragma unroll
  for (int j = 0; j < R / 2; ++j) {
    const auto pr = __builtin_amdgcn_permlane32_swap(__float_as_uint(re[2 * j]),
                                                     __float_as_uint(re[2 * j + 1]), false, false);
    const auto pi = __builtin_amdgcn_permlane32_swap(__float_as_uint(im[2 * j]),
                                                     __float_as_uint(im[2 * j + 1]), false, false);
    re[2 * j] = __uint_as_float(pr[0]);
    re[2 * j + 1] = __uint_as_float(pr[1]);
    im[2 * j] = __uint_as_float(pi[0]);
    im[2 * j + 1] = __uint_as_float(pi[1]);
  }
  dft32(re, im);  // over n1 mod 32; register q holds kappa = bitrev5(q)
}

// frequency of lane L, register q after sketch_fft: k2 + 32 k1 with
// k2 = L & 31 and k1 = 2 bitrev5(q) + (L >> 5)
__device__ __forceinline__ int fq_of(int L, int q) {
  return (L & 31) + 32 * (2 * brev5(q) + (L >> 5));
}

// One unit: (chunk of k, t, slice) = u of job j.
}  // namespace f2k
}  // namespace cm
