// CPU oracle for compressed matrix multiplication -- TEST INFRASTRUCTURE ONLY.
//
// A plain FP64 restatement of the reference algorithm (mkarppa/compmatmul).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// this library, and only to check results (or as the timed CPU baseline).
//
// Pinning (all checks live in tests/test_oracle.py):
//   orc_rng_stream     tests.cpp:90-96 known-answer vector, plus the
//                      reference's own rng.hpp compiled into oracle/_ref
//   orc_drng_stream    oracle/_ref stream; [0,1) range; 53-bit grid
//   orc_draw/orc_hash  tests.cpp:146-461 known answers and closed forms,
//                      tests.cpp:796-840 (16 sketch parameters)
//   orc_hash_arrays    the closed forms above, per entry
//   orc_fft_r2c/c2r    tests.cpp:576-615 vector, brute-force DFT, round trip
//   orc_fwht           tests.cpp:621-671 vector, brute-force Hadamard matrix,
//                      H*H = n*I
//   orc_compress       closed form orc_sketch_from_product (itself brute
//                      force over (i,j)), and the end-to-end known answers
//                      of tests.cpp:900-1003,1108-1211 and test.py
//   orc_decompress     sort-based median definition; end-to-end answers
//   orc_gemm           tests.cpp:20-88,722-750 exact products; numpy matmul
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

using cplx = std::complex<double>;

inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// rng.hpp:26-33 -- splitmix64 fills the four state words
void seed_state(uint64_t x, uint64_t s[4]) {
  for (int i = 0; i < 4; ++i) {
    x += 0x9e3779b97f4a7c15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    s[i] = z ^ (z >> 31);
  }
}

// common xoshiro256 state advance (rng.hpp:63-69, 97-103)
inline void advance(uint64_t s[4]) {
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
}

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed) { seed_state(seed, s); }
  // rng.hpp:59-72 (xoshiro256**)
  uint64_t next() {
    const uint64_t r = rotl64(s[1] * 5, 7) * 9;
    advance(s);
    return r;
  }
};

int log2_exact(uint64_t x) {
  int l = 0;
  while ((uint64_t(1) << l) < x) ++l;
  return l;
}

bool is_pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }

// iterative radix-2 complex FFT, sign -1 forward / +1 inverse, unnormalised
void fft_inplace(std::vector<cplx>& z, int sign) {
  const size_t n = z.size();
  if (n <= 1) return;
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(z[i], z[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const size_t half = len / 2;
    std::vector<cplx> w(half);
    for (size_t k = 0; k < half; ++k) {
      const double ang = 2.0 * M_PI * double(k) / double(len);
      w[k] = cplx(std::cos(ang), sign * std::sin(ang));
    }
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < half; ++k) {
        const cplx u = z[i + k];
        const cplx v = z[i + k + half] * w[k];
        z[i + k] = u + v;
        z[i + k + half] = u - v;
      }
  }
}

// real-input DFT of length n via one complex DFT of length n/2 (the
// even/odd packing of Numerical Recipes' realft); out has n/2+1 entries
void r2c(int n, const double* in, cplx* out) {
  if (n == 1) { out[0] = in[0]; return; }
  if (n == 2) { out[0] = in[0] + in[1]; out[1] = in[0] - in[1]; return; }
  const int h = n / 2;
  std::vector<cplx> z(h);
  for (int j = 0; j < h; ++j) z[j] = cplx(in[2 * j], in[2 * j + 1]);
  fft_inplace(z, -1);
  for (int k = 0; k <= h; ++k) {
    const cplx zk = z[k % h];
    const cplx zc = std::conj(z[(h - k) % h]);
    const cplx e = 0.5 * (zk + zc);
    const cplx o = cplx(0, -0.5) * (zk - zc);
    const double ang = -2.0 * M_PI * double(k) / double(n);
    out[k] = e + cplx(std::cos(ang), std::sin(ang)) * o;
  }
}

// inverse of r2c including the 1/n factor (transform.hpp:68-73)
void c2r(int n, const cplx* in, double* out) {
  if (n == 1) { out[0] = in[0].real(); return; }
  if (n == 2) {
    out[0] = 0.5 * (in[0].real() + in[1].real());
    out[1] = 0.5 * (in[0].real() - in[1].real());
    return;
  }
  const int h = n / 2;
  std::vector<cplx> z(h);
  for (int k = 0; k < h; ++k) {
    const cplx xk = in[k];
    const cplx xc = std::conj(in[h - k]);
    const cplx e = 0.5 * (xk + xc);
    const double ang = 2.0 * M_PI * double(k) / double(n);
    const cplx o = 0.5 * (xk - xc) * cplx(std::cos(ang), std::sin(ang));
    z[k] = e + cplx(0, 1) * o;
  }
  fft_inplace(z, +1);
  for (int j = 0; j < h; ++j) {
    out[2 * j] = z[j].real() / h;
    out[2 * j + 1] = z[j].imag() / h;
  }
}

void fwht_inplace(int n, double* x) {
  for (int len = 1; len < n; len <<= 1)
    for (int i = 0; i < n; i += 2 * len)
      for (int j = i; j < i + len; ++j) {
        const double a = x[j], c = x[j + len];
        x[j] = a + c;
        x[j + len] = a - c;
      }
}

inline uint32_t bucket_of(uint32_t e) { return e & 0x7fffffffu; }
inline double sign_of(uint32_t e) { return (e >> 31) ? -1.0 : 1.0; }

int nthreads(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }

}  // namespace

extern "C" {

void orc_rng_stream(uint64_t seed, int64_t n, uint64_t* out) {
  Xoshiro r(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = r.next();
}

// rng.hpp:86-111 (xoshiro256+), (r >> 11) * 2^-53
void orc_drng_stream(uint64_t seed, int64_t n, double* out) {
  uint64_t s[4];
  seed_state(seed, s);
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t r = s[0] + s[3];
    advance(s);
    out[i] = double(r >> 11) * 0x1.0p-53;
  }
}

uint64_t orc_default_seed(uint64_t millis) {
  Xoshiro r(millis);
  return r.next();
}

int64_t orc_param_words(int family) {
  switch (family) {
    case ORC_MULTIPLY_SHIFT: return 1;
    case ORC_MULTIPLY_ADD_SHIFT: return 2;
    case ORC_TABULATION8: return 4 * 256;
    case ORC_TABULATION16: return 2 * 65536;
    default: return -1;
  }
}

// hashing.hpp:69-87 (odd rejection), :145-160 (a then b), :277-295 (tables
// filled row-major with the low 32 bits of each draw)
int orc_draw(int family, int n, uint64_t buckets, uint64_t seed, uint64_t* out) {
  if (!is_pow2(int64_t(buckets))) return -1;
  const int64_t w = orc_param_words(family);
  if (w < 0) return -1;
  Xoshiro r(seed);
  for (int f = 0; f < n; ++f) {
    uint64_t* o = out + f * w;
    if (family == ORC_MULTIPLY_SHIFT) {
      uint64_t a;
      do { a = r.next(); } while ((a & 1) == 0);
      o[0] = a;
    } else if (family == ORC_MULTIPLY_ADD_SHIFT) {
      o[0] = r.next();
      o[1] = r.next();
    } else {
      for (int64_t j = 0; j < w; ++j) o[j] = uint32_t(r.next());
    }
  }
  return 0;
}

static uint32_t raw32(int family, const uint64_t* prm, uint32_t key) {
  // hashing.hpp:298-305: XOR of T[c][char c of key]
  if (family == ORC_TABULATION8) {
    uint32_t h = 0;
    for (int c = 0; c < 4; ++c) h ^= uint32_t(prm[c * 256 + ((key >> (8 * c)) & 0xff)]);
    return h;
  }
  uint32_t h = 0;
  for (int c = 0; c < 2; ++c) h ^= uint32_t(prm[c * 65536 + ((key >> (16 * c)) & 0xffff)]);
  return h;
}

// hashing.hpp:91-93, :163-165, :298-305
uint32_t orc_hash(int family, const uint64_t* prm, uint64_t buckets, uint32_t key) {
  const int lb = log2_exact(buckets);
  switch (family) {
    case ORC_MULTIPLY_SHIFT: return uint32_t((prm[0] * key) >> (64 - lb));
    case ORC_MULTIPLY_ADD_SHIFT: return uint32_t((prm[0] * key + prm[1]) >> (64 - lb));
    default: return raw32(family, prm, key) >> (32 - lb);
  }
}

// hashing.hpp:95-97, :167-169, :307-314 (top bit set -> +1)
int orc_sign(int family, const uint64_t* prm, uint32_t key) {
  switch (family) {
    case ORC_MULTIPLY_SHIFT: return ((prm[0] * key) >> 63) ? 1 : -1;
    case ORC_MULTIPLY_ADD_SHIFT: return ((prm[0] * key + prm[1]) >> 63) ? 1 : -1;
    default: return (raw32(family, prm, key) >> 31) ? 1 : -1;
  }
}

// matmul_sketch.hpp:11-20 (h1 seed, h2 seed+1, s1 seed+2, s2 seed+3)
int orc_hash_arrays(int family, int d, int b, uint64_t seed, int m, int p,
                    uint32_t* outA, uint32_t* outB) {
  const int64_t w = orc_param_words(family);
  if (w < 0 || d < 1 || !is_pow2(b) || b < 2) return -1;
  std::vector<uint64_t> h1(d * w), h2(d * w), s1(d * w), s2(d * w);
  orc_draw(family, d, b, seed, h1.data());
  orc_draw(family, d, b, seed + 1, h2.data());
  orc_draw(family, d, b, seed + 2, s1.data());
  orc_draw(family, d, b, seed + 3, s2.data());
  for (int t = 0; t < d; ++t) {
    for (int i = 0; i < m; ++i) {
      const uint32_t h = orc_hash(family, &h1[t * w], b, uint32_t(i));
      const int s = orc_sign(family, &s1[t * w], uint32_t(i));
      outA[int64_t(t) * m + i] = h | (s < 0 ? 0x80000000u : 0u);
    }
    for (int j = 0; j < p; ++j) {
      const uint32_t h = orc_hash(family, &h2[t * w], b, uint32_t(j));
      const int s = orc_sign(family, &s2[t * w], uint32_t(j));
      outB[int64_t(t) * p + j] = h | (s < 0 ? 0x80000000u : 0u);
    }
  }
  return 0;
}

int orc_fft_r2c(int n, const double* in, double* out) {
  if (!is_pow2(n)) return -1;
  r2c(n, in, reinterpret_cast<cplx*>(out));
  return 0;
}

int orc_fft_c2r(int n, const double* in, double* out) {
  if (!is_pow2(n)) return -1;
  c2r(n, reinterpret_cast<const cplx*>(in), out);
  return 0;
}

int orc_fwht(int n, double* x) {
  if (!is_pow2(n)) return -1;
  fwht_inplace(n, x);
  return 0;
}

// matmul_fft.hpp:57-122 and matmul_fwht.hpp:53-108: per inner index k and
// repetition t, count-sketch row k of AT and row k of B, transform both,
// multiply pointwise, accumulate per t; finally invert each repetition.
// Accumulation is per thread, then summed in thread order.
int orc_compress(int variant_fwht, int family, int m, int n, int p,
                 const double* AT, const double* B, int d, int b, uint64_t seed,
                 int threads, double* sketch_out) {
  if (d < 1 || b < 2 || !is_pow2(b) || m < 0 || n < 0 || p < 0) return -1;
  std::vector<uint32_t> hA(int64_t(d) * m), hB(int64_t(d) * p);
  if (orc_hash_arrays(family, d, b, seed, m, p, hA.data(), hB.data())) return -1;
  const int nt = nthreads(threads);
  const int fb = variant_fwht ? b : b / 2 + 1;  // spectrum length
  // per-thread spectral accumulators, complex for FFT, real for FWHT
  std::vector<cplx> accC(variant_fwht ? 0 : int64_t(nt) * d * fb);
  std::vector<double> accR(variant_fwht ? int64_t(nt) * d * fb : 0);
#pragma omp parallel num_threads(nt)
  {
    const int tid = omp_get_thread_num();
    std::vector<double> pa(b), pb(b);
    std::vector<cplx> fa(fb), fbv(fb);
#pragma omp for schedule(guided)
    for (int k = 0; k < n; ++k) {
      for (int t = 0; t < d; ++t) {
        std::fill(pa.begin(), pa.end(), 0.0);
        std::fill(pb.begin(), pb.end(), 0.0);
        const uint32_t* ha = &hA[int64_t(t) * m];
        const uint32_t* hb = &hB[int64_t(t) * p];
        for (int i = 0; i < m; ++i)
          pa[bucket_of(ha[i])] += sign_of(ha[i]) * AT[int64_t(k) * m + i];
        for (int j = 0; j < p; ++j)
          pb[bucket_of(hb[j])] += sign_of(hb[j]) * B[int64_t(k) * p + j];
        if (variant_fwht) {
          fwht_inplace(b, pa.data());
          fwht_inplace(b, pb.data());
          double* acc = &accR[(int64_t(tid) * d + t) * fb];
          for (int f = 0; f < b; ++f) acc[f] += pa[f] * pb[f];
        } else {
          r2c(b, pa.data(), fa.data());
          r2c(b, pb.data(), fbv.data());
          cplx* acc = &accC[(int64_t(tid) * d + t) * fb];
          for (int f = 0; f < fb; ++f) acc[f] += fa[f] * fbv[f];
        }
      }
    }
  }
  for (int t = 0; t < d; ++t) {
    double* out = sketch_out + int64_t(t) * b;
    if (variant_fwht) {
      for (int f = 0; f < b; ++f) {
        double s = 0;
        for (int th = 0; th < nt; ++th) s += accR[(int64_t(th) * d + t) * fb + f];
        out[f] = s;
      }
      fwht_inplace(b, out);
      for (int f = 0; f < b; ++f) out[f] *= 1.0 / b;
    } else {
      std::vector<cplx> sum(fb);
      for (int f = 0; f < fb; ++f) {
        cplx s = 0;
        for (int th = 0; th < nt; ++th) s += accC[(int64_t(th) * d + t) * fb + f];
        sum[f] = s;
      }
      c2r(b, sum.data(), out);
    }
  }
  return 0;
}

double orc_median(int d, double* xs) {
  std::nth_element(xs, xs + d / 2, xs + d);
  return xs[d / 2];
}

int orc_decompress(int variant_fwht, int m, int p, int d, int b,
                   const uint32_t* hA, const uint32_t* hB, const double* sketch,
                   int threads, double* C) {
  const uint32_t mask = uint32_t(b - 1);
#pragma omp parallel num_threads(nthreads(threads))
  {
    std::vector<double> xs(d);
#pragma omp for schedule(guided)
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < p; ++j) {
        for (int t = 0; t < d; ++t) {
          const uint32_t ea = hA[int64_t(t) * m + i], eb = hB[int64_t(t) * p + j];
          const uint32_t h = variant_fwht ? (bucket_of(ea) ^ bucket_of(eb))
                                          : ((bucket_of(ea) + bucket_of(eb)) & mask);
          xs[t] = sign_of(ea) * sign_of(eb) * sketch[int64_t(t) * b + h];
        }
        C[int64_t(i) * p + j] = orc_median(d, xs.data());
      }
  }
  return 0;
}

int orc_sketch_from_product(int variant_fwht, int m, int p, int d, int b,
                            const uint32_t* hA, const uint32_t* hB,
                            const double* C, double* out) {
  const uint32_t mask = uint32_t(b - 1);
  std::fill(out, out + int64_t(d) * b, 0.0);
  for (int t = 0; t < d; ++t)
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < p; ++j) {
        const uint32_t ea = hA[int64_t(t) * m + i], eb = hB[int64_t(t) * p + j];
        const uint32_t h = variant_fwht ? (bucket_of(ea) ^ bucket_of(eb))
                                        : ((bucket_of(ea) + bucket_of(eb)) & mask);
        out[int64_t(t) * b + h] += sign_of(ea) * sign_of(eb) * C[int64_t(i) * p + j];
      }
  return 0;
}

int orc_gemm(int m, int n, int p, const double* AT, const double* B, double* C,
             int threads) {
#pragma omp parallel for num_threads(nthreads(threads)) schedule(static)
  for (int i = 0; i < m; ++i) {
    double* c = C + int64_t(i) * p;
    std::fill(c, c + p, 0.0);
    for (int k = 0; k < n; ++k) {
      const double a = AT[int64_t(k) * m + i];
      const double* bk = B + int64_t(k) * p;
      for (int j = 0; j < p; ++j) c[j] += a * bk[j];
    }
  }
  return 0;
}

}  // extern "C"
