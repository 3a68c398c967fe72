// Host-side launchers of the HIP kernels (implemented in the .hip files).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace cm {

// ------------------------------------------------------------- hashing
struct HashArgs {
  int family;          // CM_* family id
  int d, lb, m, p;     // repetitions, log2(b), keys on the A and B side
  // multiply families: [4][d][words] (h1, h2, s1, s2), passed inline in
  // the kernel arguments when it fits, else in device memory
  const uint64_t* mul_params;
  int use_inline;
  uint64_t inline_params[256];
  // tabulation: device array [4][d][tab_words] of u32 table prefixes,
  // character c of a key indexes the prefix of table c (c*tab_len .. )
  const uint32_t* tables;
  int tab_len;         // entries kept per character table
  uint32_t* outA;      // d x m
  uint32_t* outB;      // d x p
};
hipError_t launch_hash(const HashArgs& a, hipStream_t s);

// ------------------------------------------------------------- compress
// Sliced fused compress: the spectrum of length b = b1*b2 is cut into b1
// slices of b2 frequencies; one workgroup owns (repetition t, slice, chunk
// of inner indices) and keeps its slice of the spectral accumulator on chip.
struct CompressArgs {
  int fwht;            // 1 = XOR convolution, 0 = cyclic convolution
  int f64;             // element type of AT/B/partials: 1 double, 0 float
  const void* AT; int64_t ld_at;
  const void* B; int64_t ld_b;
  const uint32_t* hA; const uint32_t* hB;
  int m, n, p, d, lb, lb2;
  int nchunks, kper;
  const void* tw;      // FFT: W_b^j, j < b, as (re, im) pairs of the element type
  void* partial;       // [nchunks][d][b1][b2] (x2 for complex)
};
int compress_lds_max_lb2(int fwht, int f64);
hipError_t launch_compress_lds(const CompressArgs& a, hipStream_t s);

// Register-resident single-wave FP32 FWHT compress (fast path): slices of
// 2^compress_fwht_reg_lb2() frequencies; `entries` is a scratch buffer of
// compress_fwht_reg_entries_bytes() bytes for the packed scatter entries.
int compress_fwht_reg_supported(int lb, int m, int p);
int compress_fwht_reg_lb2();
int compress_fwht_reg_slots();  // resident workgroups on the device
size_t compress_fwht_reg_entries_bytes(int d, int lb, int m, int p);
hipError_t launch_compress_fwht_reg(const CompressArgs& a, void* entries, hipStream_t s);

// Register-resident single-wave FP32 FFT compress (fast path): slices of
// 2^compress_fft_reg_lb2() frequencies; `scratch` holds the scatter
// schedules (compress_fft_reg_scratch_bytes()); `t2` is the table
// W_2048^(L * bitrev5(q)) at [q * 64 + L] (float2, 2048 entries).
int compress_fft_reg_supported(int lb, int m, int p);
int compress_fft_reg_lb2();
int compress_fft_reg_slots();
size_t compress_fft_reg_scratch_bytes(int d, int lb, int m, int p);
hipError_t launch_compress_fft_reg(const CompressArgs& a, void* scratch, const void* t2,
                                   hipStream_t s);

// Sum of the per-chunk partial spectra, then the per-slice inverse transform.
struct ReduceArgs {
  int fwht, f64, f64_reduce;
  const void* partial; int nchunks;
  int d, lb, lb2;
  const void* tw;
  void* V;             // [d][b1][b2] (x2 for complex), element type
};
hipError_t launch_reduce_inverse(const ReduceArgs& a, hipStream_t s);

// Cross-slice combination and scaling into the sketch p (d x b).
struct CombineArgs {
  int fwht, f64;
  const void* V; int d, lb, lb2;
  const void* tw;
  void* P;
};
hipError_t launch_combine(const CombineArgs& a, hipStream_t s);

// ------------------------------------------------------------- decompress
struct DecompressArgs {
  int fwht, f64;
  const void* P; const uint32_t* hA; const uint32_t* hB;
  int m, p, d, b;
  void* C; int64_t ldc;      // may be null when only recovering
  int recover; double tau;
  unsigned long long* counter;
  int64_t max_out; int32_t* oi; int32_t* oj; void* ov;
};
hipError_t launch_decompress(const DecompressArgs& a, hipStream_t s);

// ------------------------------------------------------------- plain path
// Stage-by-stage pipeline that materialises per-k sketches and spectra.
struct PlainArgs {
  int fwht, f64;
  const void* AT; int64_t ld_at;
  const void* B; int64_t ld_b;
  const uint32_t* hA; const uint32_t* hB;
  int m, n, p, d, lb;
  int t;               // repetition handled by this call
  const void* tw;
  void* SA; void* SB;  // n x b (x2 for complex) sketches / spectra
  void* acc;           // b (x2 for complex) spectrum of repetition t
};
hipError_t launch_plain_sketch(const PlainArgs& a, hipStream_t s);
hipError_t launch_plain_transform(const PlainArgs& a, hipStream_t s);
hipError_t launch_plain_mac(const PlainArgs& a, hipStream_t s);

// ------------------------------------------------------------- gemm
hipError_t launch_gemm_f64(int m, int n, int p, const double* AT, const double* B,
                           double* C, hipStream_t s);

}  // namespace cm
