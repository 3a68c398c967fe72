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
// Many products (device table of HashArgs, use_inline = 0): blocks[g] =
// (job, workgroup within the job), hash_batch_blocks() workgroups per job.
int hash_batch_blocks(const HashArgs& a);
hipError_t launch_hash_batch(const HashArgs* jobs, const int2* blocks, int nblocks, hipStream_t s);

// ------------------------------------------------------------- tables
// W_b^j (j < b) as (re, im) pairs of the element type, and the 2048-point
// register DFT's per-lane table W_2048^(L * bitrev5(q)) at [q * 64 + L]
// (float2, 2048 entries); both computed in FP64 on the device (tables.hip).
size_t twiddle_bytes(int lb, int f64);
hipError_t launch_twiddles(int lb, int f64, void* out, hipStream_t s);
hipError_t launch_t2_table(void* out, hipStream_t s);

// ------------------------------------------------------------- FFT balance
// Per inner index k, the power-of-two scales (2^e_k, 2^-e_k) applied to row
// k of AT and row k of B before pair packing, which equalise the two halves
// of z_k = S a_k + i S' b_k (both 0 when either row is zero); balance.hip.
struct PairScaleArgs {
  int f64;                  // element type of AT, B and out
  const void* AT; int64_t ld_at;
  const void* B; int64_t ld_b;
  int m, n, p;
  void* out;                // n pairs (scale of a_k, scale of b_k)
};
int pair_scale_blocks(const PairScaleArgs& a);
hipError_t launch_pair_scale(const PairScaleArgs& a, hipStream_t s);
// Many FP32 products: blocks[g] = (job, workgroup within the job).
hipError_t launch_pair_scale_batch(const PairScaleArgs* jobs, const int2* blocks, int nblocks,
                                   hipStream_t s);

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
  const void* scale;   // FFT: per k, (scale of row k of AT, of row k of B) (balance.hip)
  void* partial;       // [nchunks][d][b1][b2] (x2 for complex)
};
int compress_lds_max_lb2(int fwht, int f64);
int compress_lds_slots(int fwht, int f64, int lb2);  // resident workgroups of the general kernel
hipError_t launch_compress_lds(const CompressArgs& a, hipStream_t s);

// ------------------------------------------------- register-resident compress
// FP32 fast paths (compress_reg.hip: FWHT, compress_fft_reg.hip: FFT).  One
// 64-lane workgroup transforms 2048-frequency slices in registers; the
// spectrum of length b = b1 * 2048 is cut into b1 = 2^lbs slices.  A product
// with b < 2048 runs as one zero-padded 2048-point slice.  Padding is exact
// for both variants: the padded XOR convolution repeats the length-b one,
// and the padded cyclic convolution is the linear one, which the combine
// kernel folds back to length b.
constexpr int kRegLb2 = 11;
struct RegJob {
  const float* AT; int64_t ld_at;
  const float* B; int64_t ld_b;
  const uint32_t* hA; const uint32_t* hB;
  int m, n, p, d, lb, lbs;  // lbs = reg_lbs(lb)
  int kper, nchunks;        // inner indices per unit; units = nchunks * d << lbs
  void* lists;              // scatter schedules (*_reg_lists_bytes())
  const void* tw;           // FFT: W_b^j (float2), j < max(b, 2048)
  const void* t2;           // FFT: W_2048^(L * bitrev5(q)) at [q * 64 + L]
  const float2* scale;      // FFT: per k, (scale of row k of AT, of row k of B) (balance.hip)
  float* partial;           // [nchunks][d][b1][2048] float (float2 for FFT)
};
inline int reg_lbs(int lb) { return lb > kRegLb2 ? lb - kRegLb2 : 0; }

int fwht_reg_supported(int lb, int m, int p);
int fwht_reg_slots();  // resident workgroups of the compress kernel
size_t fwht_reg_lists_bytes(int d, int lb, int m, int p);
hipError_t launch_fwht_reg(const RegJob& j, hipStream_t s);  // schedules + compress
int fft_reg_supported(int lb, int m, int p);
int fft_reg_slots();
size_t fft_reg_lists_bytes(int d, int lb, int m, int p);
hipError_t launch_fft_reg(const RegJob& j, hipStream_t s);

// Slice-group FP32 FWHT compress (compress_grp.hip): a workgroup of
// 2^fwht_grp_log_g(lb) waves scatters each row once for that many slices.
// Same RegJob and partial layout as the single-wave kernel; lists from
// fwht_grp_lists_bytes().
int fwht_grp_supported(int lb, int m, int p);
int fwht_grp_log_g(int lb);
int fwht_grp_slots(int lb);  // resident workgroups
size_t fwht_grp_lists_bytes(int d, int lb, int m, int p);
hipError_t launch_fwht_grp(const RegJob& j, hipStream_t s);

// Slice-group FP32 FFT compress (compress_fft_grp.hip): a workgroup of
// 2^fft_grp_log_g(lb) waves owns the strided slices s0 + j b1/G and
// scatters each pair-packed row once for all of them.
int fft_grp_supported(int lb, int m, int p);
int fft_grp_log_g(int lb);
int fft_grp_slots(int lb);
size_t fft_grp_lists_bytes(int d, int lb, int m, int p);
hipError_t launch_fft_grp(const RegJob& j, hipStream_t s);
// Many products on the slice-group kernels (workgroups of 2^log_g waves):
// schedules and the persistent compress launch, as for the single-wave ones.
hipError_t launch_fwht_grp_lists_batch(int log_g, const RegJob* jobs, const int2* blocks, int nblocks,
                                       hipStream_t s);
hipError_t launch_fwht_grp_batch(int log_g, const RegJob* jobs, const int2* items, const int* qstart,
                                 unsigned* counters, int grid, hipStream_t s);
hipError_t launch_fft_grp_lists_batch(int log_g, const RegJob* jobs, const int2* blocks, int nblocks,
                                      hipStream_t s);
hipError_t launch_fft_grp_batch(int log_g, const RegJob* jobs, const int2* items, const int* qstart,
                                unsigned* counters, int grid, hipStream_t s);

// Many products.  Schedules: `blocks[g]` = (job, list) of workgroup g.
// Compress: one persistent launch drains eight work queues, queue x holding
// items[qstart[x] .. qstart[x + 1]) = (job, unit); the workgroups of XCD x
// take queue x first, then the others.  counters[0..8) must be zero.
hipError_t launch_fwht_reg_lists_batch(const RegJob* jobs, const int2* blocks, int nblocks,
                                       hipStream_t s);
hipError_t launch_fwht_reg_batch(const RegJob* jobs, const int2* items, const int* qstart,
                                 unsigned* counters, int grid, hipStream_t s);
hipError_t launch_fft_reg_lists_batch(const RegJob* jobs, const int2* blocks, int nblocks,
                                      hipStream_t s);
hipError_t launch_fft_reg_batch(const RegJob* jobs, const int2* items, const int* qstart,
                                unsigned* counters, int grid, hipStream_t s);
// lists per job for the schedule launch (2 per (t, slice) for FWHT: A and B)
inline int reg_lists_count(int fwht, int d, int lb) { return (fwht ? 2 : 1) * (d << reg_lbs(lb)); }
hipError_t launch_zero_u32(unsigned* p, int n, hipStream_t s);

// Matrix-core FFT compress (fft_mfma.hip): radix-16 Stockham stages as
// hi/lo-split BF16 or FP16 MFMA products, b = 16^S with 2^12 <= b <= 2^28.
// Writes the per-chunk spectra in the sliced layout the reduce kernel reads
// (slices of 2^kRegLb2).
struct Bf16FftArgs {
  const float* AT; int64_t ld_at;
  const float* B; int64_t ld_b;
  const uint32_t* hA; const uint32_t* hB;
  int m, n, p, d, lb;
  int rows, nchunks;       // inner indices per chunk, chunks
  const void* tw;          // W_b^j (float2)
  const float2* scale;     // per k, (scale of row k of AT, of row k of B) (balance.hip)
  float2* Z0; float2* Z1;  // rows x b complex each
  float* partial;          // [nchunks][d][b1][2048] complex
  int fp16;                // 1: FP16 operands (column-scaled), 0: BF16
};
int fft_bf16_supported(int lb);
hipError_t launch_fft_bf16_compress(const Bf16FftArgs& a, hipStream_t s);

// Sum of the per-chunk partial spectra, then the per-slice inverse transform.
struct ReduceArgs {
  int fwht, f64, f64_reduce;
  const void* partial; int nchunks;
  int d, lb, lb2;
  const void* tw;
  void* V;             // [d][b1][b2] (x2 for complex), element type
  // one slice (lb2 = lb) and no padding fold: the sketch p (d x b), written
  // by the inverse itself with the combine kernel's arithmetic (no combine)
  void* P = nullptr;
};
hipError_t launch_reduce_inverse(const ReduceArgs& a, hipStream_t s);

// Cross-slice combination and scaling into the sketch p (d x 2^lb_out).
// lb_out < lb folds a zero-padded product (lb = 11) back to its length.
struct CombineArgs {
  int fwht, f64;
  const void* V; int d, lb, lb2;
  const void* tw;
  void* P;
  int lb_out;
};
hipError_t launch_combine(const CombineArgs& a, hipStream_t s);

// Many FP32 products: chunk sum, per-slice inverse and combine of every job
// in one launch each; `*_blocks[g]` = (job, workgroup within the job), with
// the per-job counts from finish_batch_blocks() (returns the combine count).
struct FinishJob { ReduceArgs r; CombineArgs c; };
int finish_batch_blocks(const FinishJob& j, int* sum_blocks, int* red_blocks);
// LDS bytes of one reduce workgroup: the slice in a tile padded by one
// element every 16 (complex for FFT, rounded to 16 bytes), and for FFT the
// two-level twiddle tables of its length (64 + b2 / 64 + 1 entries).
inline size_t reduce_lds_bytes(bool fwht, size_t real_bytes, int lb2) {
  const size_t b2 = size_t(1) << lb2;
  const size_t elem = fwht ? real_bytes : 2 * real_bytes;
  const size_t tile = ((b2 + b2 / 16) * elem + 15) & ~size_t(15);
  return fwht ? tile : tile + (64 + b2 / 64 + 1) * elem;
}

hipError_t launch_finish_batch(const FinishJob* jobs, const int2* sum_blocks, int n_sum,
                               const int2* red_blocks, int n_red, size_t red_lds,
                               const int2* comb_blocks, int n_comb, hipStream_t s);

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
// Many FP32 products of one class (variant, d <= 16, form): blocks[g] =
// (job, workgroup within the job), with decompress_batch_blocks() workgroups
// per job.  Forms, by the sketch row's length (decompress_batch_form):
// 0 gathers from L2 (rows above 64 KB, or d > 16), 1 two LDS buffers (rows up
// to 16 KB), 2 the ring of five 32 KB LDS slots for 32 KB rows, 3 the ring
// for 64 KB rows (half rows).  The ring forms need their 160 KB LDS
// attribute, which decompress_batch_form sets (at plan time, outside any
// graph capture); it returns a gathering form where the attribute is refused.
// `lds` = decompress_batch_lds() of the class's largest row.
int decompress_batch_form(int fwht, int d, int b);
size_t decompress_batch_lds(int form, int b);
int decompress_batch_blocks(int d, int form, int m, int p);
hipError_t launch_decompress_batch(int fwht, int d, int form, size_t lds,
                                   const DecompressArgs* jobs, const int2* blocks, int nblocks,
                                   hipStream_t s);

// ------------------------------------------------------------- plain path
// Stage-by-stage pipeline that materialises per-k sketches and spectra in
// HBM (plain.hip).  One pass covers repetitions [t0, t0 + nt) and inner
// indices [k0, k0 + nk) (either nt = 1 or the whole of k): nt * nk rows of b
// elements in SA (FFT: complex, pair-packed) and SB (FWHT: B's sketches).
struct PlainArgs {
  int fwht, f64;
  const void* AT; int64_t ld_at;
  const void* B; int64_t ld_b;
  const uint32_t* hA; const uint32_t* hB;
  int m, n, p, d, lb;
  int t0, nt, k0, nk;
  const void* tw;      // W_b^j, j < b, element type
  const void* scale;   // FFT: per k, (scale of row k of AT, of row k of B) (balance.hip)
  void* SA; void* SB;
  int kchunks;         // chunks of k summed separately by the MAC
  void* part;          // kchunks x nt x b partial sums (complex for FFT)
  void* acc;           // d x b1 x b2 spectral sums in the reduce kernel's layout
  int lb2;             // log2 of that layout's slice length
  int first;           // this pass starts its repetitions' sums (else adds to them)
};
int plain_max_lb(int fwht, int f64);  // largest log2(b) of the plain path
hipError_t launch_plain_sketch(const PlainArgs& a, hipStream_t s);
hipError_t launch_plain_transform(const PlainArgs& a, hipStream_t s);
hipError_t launch_plain_mac(const PlainArgs& a, hipStream_t s);

// ------------------------------------------------------------- sparse operands
// CSR (rows x cols, indptr[rows + 1], indices, values) to dense row-major.
hipError_t launch_csr_dense(const int32_t* indptr, const int32_t* indices, const float* values,
                            int rows, int cols, int64_t ld, float* out, hipStream_t s);

// ------------------------------------------------------------- no-allocation scratch
// noalloc.hip: the fixed device array of the no-allocation entry point.
// noalloc_setup() looks up its address on `device` (the current device)
// once; noalloc_scratch() returns that address and the array's size.
int noalloc_setup(int device);
int noalloc_scratch(int device, void** ptr, size_t* bytes);

// ------------------------------------------------------------- gemm
hipError_t launch_gemm_f64(int m, int n, int p, const double* AT, const double* B,
                           double* C, hipStream_t s);

}  // namespace cm
