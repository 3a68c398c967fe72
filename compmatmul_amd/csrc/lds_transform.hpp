// LDS-resident radix-2^r transforms shared by the plain pipeline and the
// general (FP64) fused kernel: register DFT/WHT kernels of up to 16 points,
// decimation-in-frequency passes over padded LDS tiles with two-level LDS
// twiddle tables, and the digit order of the FFT output.  Radix 2 / 4 / 8 /
// 16 Cooley-Tukey and the Sylvester Hadamard butterfly, from their textbook
// definitions.
#pragma once
#include <type_traits>

#include "device_common.hpp"

namespace cm {
namespace ldsx {

template <typename T, bool FWHT>
using Elem = std::conditional_t<FWHT, T, cpx<T>>;

// one padding slot every 16 elements keeps strided passes off one bank
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }



// cos(2 pi e / 16)
__device__ __forceinline__ constexpr double cos16(int e) {
  constexpr double c[16] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508978,
                            0.0, -0.38268343236508978, -0.70710678118654752, -0.92387953251128674,
                            -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508978,
                            0.0, 0.38268343236508978, 0.70710678118654752, 0.92387953251128674};
  return c[e & 15];
}

template <int R>
constexpr int ilog2c() { return R <= 1 ? 0 : 1 + ilog2c<R / 2>(); }

// reverse the low r bits of k (compile-time)
__device__ __forceinline__ constexpr int brev_c(int k, int r) {
  int br = 0;
  for (int bit = 0; bit < r; ++bit) br |= ((k >> bit) & 1) << (r - 1 - bit);
  return br;
}

// R-point DFT (W = exp(-2 pi i / R)), natural order in and out: radix-2
// decimation in frequency in registers, then the bit-reversal undone.  Every
// loop here and in wht_reg has a constant trip count, so it unrolls
// completely and the arrays stay in registers.
template <int R, typename T>
__device__ __forceinline__ void dft_reg(cpx<T> (&v)[R]) {
  constexpr int r = ilog2c<R>();
#pragma unroll
  for (int st = 0; st < r; ++st) {
    const int len = R >> (st + 1);
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {  // butterfly q of the stage
      const int pos = q % len, j = (q / len) * 2 * len + pos;
      const cpx<T> x = v[j], y = v[j + len];
      v[j] = cadd(x, y);
      const cpx<T> dd = csub(x, y);
      const int e = pos * 8 / len;  // W_R^(pos R / 2len) = W_16^e
      if (e == 0) {
        v[j + len] = dd;
      } else if (e == 4) {  // -i
        v[j + len] = cpx<T>{dd.y, -dd.x};
      } else {
        const T c = T(cos16(e)), sn = T(cos16(e + 12));  // sin = cos(angle - pi/2)
        v[j + len] = cpx<T>{dd.x * c + dd.y * sn, dd.y * c - dd.x * sn};
      }
    }
  }
  cpx<T> o[R];
#pragma unroll
  for (int k = 0; k < R; ++k) o[k] = v[brev_c(k, r)];
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = o[k];
}

// R-point natural-order Walsh-Hadamard transform in registers
template <int R, typename T>
__device__ __forceinline__ void wht_reg(T (&v)[R]) {
  constexpr int r = ilog2c<R>();
#pragma unroll
  for (int st = 0; st < r; ++st) {
    const int len = R >> (st + 1);
#pragma unroll
    for (int q = 0; q < R / 2; ++q) {
      const int j = (q / len) * 2 * len + q % len;
      const T x = v[j], y = v[j + len];
      v[j] = x + y;
      v[j + len] = x - y;
    }
  }
}

// Twiddles of a length-2^l transform in LDS: W_L^x = hi[x >> 6] lo[x & 63]
// (lo[i] = W_L^i, hi[i] = W_L^(64 i)), filled from the length-2^lb table.
constexpr int TW_LO = 64;
template <int NT, typename T>
__device__ __forceinline__ void fill_twiddles(cpx<T>* lo, cpx<T>* hi, int l, const cpx<T>* __restrict__ tw,
                                              int lb) {
  const int L = 1 << l;
  for (int i = threadIdx.x; i < TW_LO + (L >> 6) + 1; i += NT) {
    if (i < TW_LO)
      lo[i] = i < L ? tw[int64_t(i) << (lb - l)] : cpx<T>{T(1), T(0)};
    else
      hi[i - TW_LO] = tw[(int64_t(i - TW_LO) << 6 << (lb - l)) & ((int64_t(1) << lb) - 1)];
  }
}
template <typename T>
__device__ __forceinline__ cpx<T> twiddle(const cpx<T>* lo, const cpx<T>* hi, int x) {
  return x < TW_LO ? lo[x] : cmul(hi[x >> 6], lo[x & (TW_LO - 1)]);
}

// One radix-R decimation-in-frequency pass at stride S = 2^s over `count`
// sequences of length 2^l held in LDS (sequence c from element c * cs): the
// R elements base + j S of each group are transformed, and for the FFT
// output k is multiplied by W_(R S)^(pos k) = W_L^(pos k L / (R S)), from the
// LDS twiddle tables (lo, hi).
template <int NT, typename T, bool FWHT, int R>
__device__ __forceinline__ void lds_pass(Elem<T, FWHT>* z, int l, int s, int count, int cs,
                                         const cpx<T>* lo, const cpx<T>* hi) {
  constexpr int r = ilog2c<R>();
  const int S = 1 << s;
  const int gl = l - r;  // groups per sequence, log2
  const int total = count << gl;
  for (int g = threadIdx.x; g < total; g += NT) {
    const int c = g >> gl;
    const int gg = g & ((1 << gl) - 1);
    const int pos = gg & (S - 1);
    const int base = c * cs + ((gg >> s) << (s + r)) + pos;
    Elem<T, FWHT> v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) v[j] = z[pad(base + j * S)];
    if constexpr (FWHT) {
      wht_reg<R>(v);
    } else {
      dft_reg<R>(v);
      if (pos != 0) {
        const int sh = l - s - r;
        if constexpr (sizeof(T) == 8) {
          // FP64: W^(pos k) by recurrence from one table lookup (at most R - 2
          // products, about 2e-15 relative), sparing 2 (R - 2) LDS reads
          const cpx<T> w1 = twiddle(lo, hi, pos << sh);
          cpx<T> w = w1;
#pragma unroll
          for (int k = 1; k < R; ++k) {
            v[k] = cmul(v[k], w);
            if (k + 1 < R) w = cmul(w, w1);
          }
        } else {
#pragma unroll
          for (int k = 1; k < R; ++k) v[k] = cmul(v[k], twiddle(lo, hi, (pos * k) << sh));
        }
      }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) z[pad(base + j * S)] = v[j];
  }
}

// widest pass, log2: radix 16, or 8 for the FP64 FFT (whose radix-16 pass
// would spill at 1024 threads)
template <typename T, bool FWHT>
constexpr int rmax() { return !FWHT && sizeof(T) == 8 ? 3 : 4; }

// Transform `count` LDS sequences of length 2^l with NT threads: radix-2^RM
// passes (RM = 4 unless given) from the high bits down, the last pass over
// the l mod RM lowest bits (if any).  FWHT output is in natural order; FFT
// output k sits at dif_pos(k, l, RM).
template <int NT, typename T, bool FWHT, int RM = rmax<T, FWHT>()>
__device__ __forceinline__ void lds_transform(Elem<T, FWHT>* z, int l, int count, int cs, const cpx<T>* lo,
                              const cpx<T>* hi) {
  int s = l;
  while (s > 0) {
    const int r = min(RM, s);
    s -= r;
    __syncthreads();
    switch (r) {
      case 4:
        if constexpr (RM >= 4) lds_pass<NT, T, FWHT, 16>(z, l, s, count, cs, lo, hi);
        break;
      case 3: lds_pass<NT, T, FWHT, 8>(z, l, s, count, cs, lo, hi); break;
      case 2: lds_pass<NT, T, FWHT, 4>(z, l, s, count, cs, lo, hi); break;
      default: lds_pass<NT, T, FWHT, 2>(z, l, s, count, cs, lo, hi); break;
    }
  }
  __syncthreads();
}

// log2 of the radix of lds_transform's last pass (the lowest bits)
__host__ __device__ __forceinline__ constexpr int last_radix_log2(int l, int rm) {
  return l > 0 ? (l - 1) % rm + 1 : 0;
}

// lds_transform without its last pass: the passes over the bits above the
// lowest last_radix_log2(l, RM), each after a barrier, then a barrier.  The
// last pass (stride 1, no twiddles) transforms the R = 2^last_radix_log2
// elements g R + j of each group g; a caller that consumes the transform
// right away runs it in registers (last_pass) instead of through LDS.
template <int NT, typename T, bool FWHT, int RM = rmax<T, FWHT>()>
__device__ __forceinline__ void lds_transform_head(Elem<T, FWHT>* z, int l, int count, int cs,
                                                   const cpx<T>* lo, const cpx<T>* hi) {
  const int rl = last_radix_log2(l, RM);
  for (int s = l - RM; s >= rl; s -= RM) {
    __syncthreads();
    lds_pass<NT, T, FWHT, (1 << RM)>(z, l, s, count, cs, lo, hi);
  }
  __syncthreads();
}

// the last pass of lds_transform on one group's R elements, in registers
template <int R, typename T, bool FWHT>
__device__ __forceinline__ void last_pass(Elem<T, FWHT> (&v)[R]) {
  if constexpr (FWHT)
    wht_reg<R>(v);
  else
    dft_reg<R>(v);
}

// LDS of a tile of 2^l elements (padded), rounded to 16 bytes, then (FFT)
// the twiddle tables lo (64 entries) and hi (2^l / 64 + 1)
template <typename T, bool FWHT>
__host__ __device__ __forceinline__ size_t tile_bytes(int l) {
  return ((size_t((1 << l) + (1 << l) / 16) * sizeof(Elem<T, FWHT>)) + 15) & ~size_t(15);
}
template <typename T, bool FWHT>
__host__ __device__ __forceinline__ size_t transform_bytes(int l) {
  return tile_bytes<T, FWHT>(l) + (FWHT ? 0 : (64 + (1 << l) / 64 + 1) * sizeof(cpx<T>));
}
template <typename T, bool FWHT>
__device__ __forceinline__ cpx<T>* twiddle_lds(unsigned char* smem, int l) {
  return reinterpret_cast<cpx<T>*>(smem + tile_bytes<T, FWHT>(l));
}

// 16-byte vectors of elements for the row loads and stores
template <typename E>
struct alignas(16) Vec16 {
  E e[16 / sizeof(E)];
};

// Position of DFT output k after lds_transform with passes of rm bits: the
// radix digits of k (rm bits each from the low end, the last one l mod rm
// bits) in reverse order.
__device__ __forceinline__ int dif_pos(int k, int l, int rm) {
  int p = 0, off = l;
  while (off > 0) {
    const int r = min(rm, off);
    off -= r;
    p |= (k & ((1 << r) - 1)) << off;
    k >>= r;
  }
  return p;
}


// Natural index of the FFT output at position p (the inverse of dif_pos).
__device__ __forceinline__ int dif_nat(int p, int l, int rm) {
  int k = 0, off = l, sh = 0;
  while (off > 0) {
    const int r = min(rm, off);
    off -= r;
    k |= ((p >> off) & ((1 << r) - 1)) << sh;
    sh += r;
  }
  return k;
}

}  // namespace ldsx
}  // namespace cm
