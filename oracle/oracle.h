/*
 * CPU oracle for compressed matrix multiplication -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain FP64 C++ restatement of the reference algorithm
 * (Pagh's compressed matrix multiplication as implemented in rng.hpp,
 * hashing.hpp, matmul_sketch.hpp, transform.hpp, matmul_fft.hpp,
 * matmul_fwht.hpp and array.hpp of mkarppa/compmatmul).  It is used only to
 * check results, by the tests and by __graft_entry__.smoke(), and as the
 * cpu_baseline leg of bench.py.  Nothing in the product library links,
 * loads or calls it.
 *
 * Every function names the reference lines it follows.  How each function
 * is pinned (known-answer vectors, closed forms, brute force) is listed at
 * the definition in oracle.cpp; every function has at least one such check
 * in tests/test_oracle.py.
 */
#ifndef CM_ORACLE_H
#define CM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* hash family ids (same numbering as the product C-ABI) */
enum { ORC_MULTIPLY_ADD_SHIFT = 0, ORC_MULTIPLY_SHIFT = 1,
       ORC_TABULATION8 = 2, ORC_TABULATION16 = 3 };

/* rng.hpp:26-33,59-72: splitmix64-seeded xoshiro256**, n outputs */
void orc_rng_stream(uint64_t seed, int64_t n, uint64_t* out);
/* rng.hpp:86-111: splitmix64-seeded xoshiro256+, doubles in [0,1) */
void orc_drng_stream(uint64_t seed, int64_t n, double* out);
/* rng.hpp:40-44 + matmul_fft.hpp:30-32: seed used when none is given */
uint64_t orc_default_seed(uint64_t millis);

/* Number of uint64 parameter words one hash function of `family` holds:
 * 1 (multiply-shift a), 2 (multiply-add-shift a,b), 1024 (tab8 T[4][256])
 * or 131072 (tab16 T[2][65536]); tabulation words hold uint32 values. */
int64_t orc_param_words(int family);
/* hashing.hpp draw(n, buckets, seed): parameters of n functions drawn from
 * one rng stream; out has n*orc_param_words(family) words. */
int orc_draw(int family, int n, uint64_t buckets, uint64_t seed, uint64_t* out);
/* hashing.hpp operator()/sign for one function with parameters `params` */
uint32_t orc_hash(int family, const uint64_t* params, uint64_t buckets, uint32_t key);
int orc_sign(int family, const uint64_t* params, uint32_t key);

/* Bucket/sign arrays of a sketch (matmul_sketch.hpp:11-20): h1 from seed,
 * h2 from seed+1, s1 from seed+2, s2 from seed+3.  Packed per entry as
 * bucket | (sign < 0 ? 1u<<31 : 0).  outA is d x m (t-major), outB d x p. */
int orc_hash_arrays(int family, int d, int b, uint64_t seed, int m, int p,
                    uint32_t* outA, uint32_t* outB);

/* transform.hpp:57-61 (FFTW r2c, e^{-2 pi i kn/n}, unnormalised):
 * in n doubles -> out (n/2+1) complex as interleaved re,im. */
int orc_fft_r2c(int n, const double* in, double* out);
/* transform.hpp:68-73: c2r then x 1/n */
int orc_fft_c2r(int n, const double* in, double* out);
/* transform.hpp:110-112 (walsh_wak: natural order, unnormalised), in place */
int orc_fwht(int n, double* x);

/* matmul_fft.hpp:25-129 / matmul_fwht.hpp:25-114: the compressed product.
 * AT is n x m row-major, B is n x p row-major; sketch_out is d x b (FP64).
 * threads <= 0 means all OpenMP threads. */
int orc_compress(int variant_fwht, int family, int m, int n, int p,
                 const double* AT, const double* B, int d, int b, uint64_t seed,
                 int threads, double* sketch_out);
/* matmul_fft.hpp:135-147,165-177 / matmul_fwht.hpp:122-134,154-166:
 * C[i,j] = median_t(s1t(i) s2t(j) p[t, h]) with h = (h1+h2)&(b-1) or h1^h2 */
int orc_decompress(int variant_fwht, int m, int p, int d, int b,
                   const uint32_t* hA, const uint32_t* hB, const double* sketch,
                   int threads, double* C);
/* array.hpp:323-328: nth_element at d/2 (upper median for even d) */
double orc_median(int d, double* xs);
/* Closed form of the sketch from the dense product:
 * p_t[h] = sum over (i,j) with h1t(i) (+ or ^) h2t(j) == h of s1t(i)s2t(j)C[i,j] */
int orc_sketch_from_product(int variant_fwht, int m, int p, int d, int b,
                            const uint32_t* hA, const uint32_t* hB,
                            const double* C, double* sketch_out);
/* array.hpp:81-95: C = AT^T B in FP64 (C m x p) */
int orc_gemm(int m, int n, int p, const double* AT, const double* B, double* C,
             int threads);

#ifdef __cplusplus
}
#endif
#endif
