/*
 * compmatmul_amd -- C-ABI of the MI355X-native compressed matrix product.
 *
 * Plain pointers and sizes only; no exceptions cross this boundary.  Every
 * function returns CM_OK (0) or a negative status, and cm_last_error()
 * gives the message of the last failure on the calling thread.
 *
 * Notation follows the reference: m = n_a (rows of C), n = n_inner,
 * p = n_b (columns of C).  AT is n x m row-major (A passed transposed),
 * B is n x p row-major, C is m x p row-major (compmatmul.cpp:55-63).
 * Calls without a context argument run on the library's default context
 * (one HIP stream on device 0, or cm_init's device); contexts (below) have
 * their own device, stream and memory.  Calls are synchronous with respect
 * to the host unless stated otherwise.
 */
#ifndef COMPMATMUL_AMD_H
#define COMPMATMUL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CM_OK 0
#define CM_EINVAL (-1)    /* invalid argument (ValueError at the Python layer) */
#define CM_ENOMEM (-2)    /* device or host allocation failed (MemoryError) */
#define CM_EDEVICE (-3)   /* HIP runtime error or no usable gfx950 device */
#define CM_ENOTSUP (-4)   /* argument combination not supported */
#define CM_EHASH (-5)     /* unknown hash family (RuntimeError) */

/* hash families; the numbering is this library's own */
#define CM_MULTIPLY_ADD_SHIFT 0 /* hashing.hpp:111-183, the default */
#define CM_MULTIPLY_SHIFT 1     /* hashing.hpp:27-102 */
#define CM_TABULATION8 2        /* hashing.hpp:232-320, 8-bit characters */
#define CM_TABULATION16 3       /* hashing.hpp:232-320, 16-bit characters */

/* transform variant */
#define CM_FFT 0  /* cyclic convolution, matmul_fft.hpp */
#define CM_FWHT 1 /* XOR convolution, matmul_fwht.hpp */

/* element type of AT, B, C and of the sketch arithmetic */
#define CM_F64 0
#define CM_F32 1

/* where a pointer lives */
#define CM_HOST 0
#define CM_DEVICE 1

/* flags for cm_compress_ex */
#define CM_FLAG_F64_REDUCE 1u /* FP32 path: sum per-tile spectra in FP64 */
#define CM_FLAG_PLAIN 2u      /* unfused stage-by-stage path (sketches and
                                 spectra materialised in HBM) */
#define CM_FLAG_LDS_KERNEL 4u /* FP32: use the general LDS-resident fused
                                 kernel instead of the register-resident ones */
#define CM_FLAG_SINGLE_WAVE 16u /* FP32: one wave per slice instead of the
                                   slice-group kernels (comparison) */
#define CM_FLAG_BF16_MFMA 8u  /* FP32 FFT, b = 16^k (2^12 .. 2^28): forward
                                 transforms as radix-16 DFT stages on BF16
                                 matrix cores with hi/lo split correction and
                                 FP32 accumulation (precision study mode) */
#define CM_FLAG_FP16_MFMA 32u /* as CM_FLAG_BF16_MFMA with FP16 operands:
                                 each DFT-16 input column scaled by a power
                                 of two into FP16's range before the split */

typedef struct cm_sketch cm_sketch; /* device-resident compressed product */

/* Message of the last failing call on this thread ("" if none). */
const char* cm_last_error(void);

/* Select the HIP device used by later calls (default 0) and create its
 * stream.  Optional: every entry point initialises device 0 on first use. */
int cm_init(int device);

/* Seed used when the caller gives none: the first xoshiro256** output of a
 * generator seeded with the wall-clock milliseconds.
 * Replaces: matmul_fft.hpp:30-32 / matmul_fwht.hpp:30-32 (rng.hpp:40-57). */
uint64_t cm_default_seed(void);

/* ---------------------------------------------------------------- drop-in
 * C = A^T B estimated by the compressed product, FP64, host pointers,
 * synchronous; C is fully overwritten.  seed_given = 0 draws cm_default_seed().
 * Replaces: compmatmul::matmul_fft<Hash>(n_a, n_inner, n_b, C, AT, B, d, b,
 * seed) (matmul_fft.hpp:155-178), dispatched by compmatmul.cpp:107-146. */
int cm_matmul_fft(int n_a, int n_inner, int n_b, double* C, const double* AT,
                  const double* B, int d, int b, uint64_t seed, int seed_given,
                  int family);

/* Replaces: compmatmul::matmul_fwht<Hash> (matmul_fwht.hpp:144-167),
 * dispatched by compmatmul.cpp:149-185. */
int cm_matmul_fwht(int n_a, int n_inner, int n_b, double* C, const double* AT,
                   const double* B, int d, int b, uint64_t seed, int seed_given,
                   int family);

/* Exact FP64 product C = AT^T B on the device, host pointers.
 * Replaces: compmatmul::array::matmul_gemm<double> (array.hpp:81-95) behind
 * compmatmul.cpp:76-85. */
int cm_gemm(int n_a, int n_inner, int n_b, double* C, const double* AT,
            const double* B);

/* ---------------------------------------------------------- no allocation
 * Size of the fixed device scratch behind cm_matmul_noalloc (per GPU, held
 * until the process ends). */
#define CM_NOALLOC_SCRATCH_BYTES (64u << 20)

/* C = the median-of-d estimate of AT^T B, as cm_matmul_fft / cm_matmul_fwht
 * compute it, for device operands: AT (n x m, leading dimension ld_at), B
 * (n x p, ld_b) and C (m x p, ldc) are device pointers of dtype CM_F32 or
 * CM_F64.  Every intermediate (hash arrays, scatter schedules, partial
 * spectra, sketch, twiddle tables) lives in a fixed device array of
 * CM_NOALLOC_SCRATCH_BYTES, so the call allocates no device or host memory
 * and uploads nothing; it returns CM_ENOMEM, naming the size needed, when a
 * product's intermediates do not fit, and CM_ENOTSUP for the tabulation
 * families or more hash parameters than travel in kernel arguments (d > 32
 * for multiply-add-shift, d > 64 for multiply-shift).  variant CM_FFT or
 * CM_FWHT; h1, h2, s1, s2 from seed + 0..3.  Synchronous: C is complete on
 * return.  Every call on a GPU shares that GPU's scratch (it keeps the last
 * call's contents); calls through the library are serialised by its lock,
 * and calls that overlapped on one GPU would race on the scratch.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). */
int cm_matmul_noalloc(int variant, int dtype, int n_a, int n_inner, int n_b, const void* AT,
                      int64_t ld_at, const void* B, int64_t ld_b, void* C, int64_t ldc, int d,
                      int b, uint64_t seed, int family);

/* ---------------------------------------------------------------- hashing
 * Packed bucket/sign arrays of a sketch: outA[t*m+i] = h1_t(i) | signbit,
 * outB[t*p+j] = h2_t(j) | signbit, where signbit = 1u<<31 when s1_t(i)
 * (resp. s2_t(j)) is -1.  Functions h1,h2,s1,s2 are drawn from seed,
 * seed+1, seed+2, seed+3.  Evaluated on the device; `mem` says where
 * outA/outB live.
 * Replaces: matmul_sketch<Hash>(d, b, seed) (matmul_sketch.hpp:11-20) and
 * the per-key Hash::operator()/sign (hashing.hpp:91-97, 163-169, 298-314). */
int cm_hash_arrays(int family, int d, int b, uint64_t seed, int m, int p,
                   uint32_t* outA, uint32_t* outB, int mem);

/* ---------------------------------------------------------------- sketches
 * Compressed product of AT (n x m, leading dimension ld_at) and B (n x p,
 * leading dimension ld_b) into a device-resident sketch of d x b values.
 * dtype CM_F64 or CM_F32 (inputs and arithmetic); mem CM_HOST or CM_DEVICE.
 * Host operands are read before the call returns (through pinned staging),
 * so their memory may be reused at once; device operands are read on the
 * stream.  *out receives a new handle (free with cm_sketch_free).
 * Replaces: matmul_compressed_product_fft<Hash> (matmul_fft.hpp:25-129) and
 * matmul_compressed_product_fwht<Hash> (matmul_fwht.hpp:25-114). */
int cm_compress(int variant, int dtype, int n_a, int n_inner, int n_b,
                const void* AT, int64_t ld_at, const void* B, int64_t ld_b,
                int d, int b, uint64_t seed, int family, int mem, cm_sketch** out);

/* As cm_compress, reusing an existing handle when its (variant, dtype, d,
 * b, m, p) match (no device allocation after the first call), with flags. */
int cm_compress_ex(int variant, int dtype, int n_a, int n_inner, int n_b,
                   const void* AT, int64_t ld_at, const void* B, int64_t ld_b,
                   int d, int b, uint64_t seed, int family, int mem,
                   unsigned flags, cm_sketch** inout);

/* Full median-of-d decompression into C (m x p, leading dimension ldc, the
 * sketch's dtype).  Replaces: the decompress loops of matmul_fft.hpp:165-177
 * and matmul_fwht.hpp:154-166 (matmul_decompress_fft/_fwht, array::median). */
int cm_decompress(const cm_sketch* sk, void* C, int64_t ldc, int mem);

/* Large-entry recovery: every (i, j) whose median estimate is > tau, sorted
 * by (i, j), written to host arrays of capacity max_out; *count receives the
 * total number found (which may exceed max_out).  No counterpart function in
 * the reference, which leaves thresholding of the full C to the caller. */
int cm_recover(const cm_sketch* sk, double tau, int64_t max_out, int32_t* out_i,
               int32_t* out_j, double* out_val, int64_t* count);

/* Copy the d x b sketch values (as FP64) to host memory.
 * Replaces: reading matmul_sketch<Hash>::p (matmul_sketch.hpp:80). */
int cm_sketch_values(const cm_sketch* sk, double* host_out);

/* Copy the packed hash arrays of a sketch to host memory (d*m and d*p). */
int cm_sketch_hashes(const cm_sketch* sk, uint32_t* outA, uint32_t* outB);

void cm_sketch_free(cm_sketch* sk);

/* ---------------------------------------------------------------- contexts
 * A context holds a device, a stream and the working memory of the products
 * run through it; the entry points without a context argument use the
 * library's default context (device 0 on first use, or cm_init's).  Calls on
 * one context are serialized by its own lock, and calls on different
 * contexts run concurrently (from different threads, on their streams).
 * The reference has no such object: each compmatmul call allocates its
 * sketch and per-thread scratch and frees them on return
 * (compmatmul.cpp:107-147, matmul_fft.hpp:57-70, matmul_fwht.hpp:53-63);
 * a context keeps that state on the device between calls. */
typedef struct cm_context cm_context;

/* device: HIP device index.  stream: the caller's hipStream_t, or NULL for
 * a non-blocking stream the context creates and destroys.  workspace_bytes:
 * 0 = per-call buffers grow on demand, as the default context's do; > 0 =
 * one device region of that size, allocated here, from which every per-call
 * intermediate (staged host operands, hash parameters and tables, balance
 * scales, twiddle tables, scatter schedules, partial spectra, inverse,
 * decompression and recovery buffers) is carved; a call whose
 * intermediates do not fit fails with CM_ENOMEM, naming the size it needs,
 * and the plain path and the matrix-core modes (which allocate their own
 * buffers) fail with CM_ENOTSUP.  Sketches are owned by their handles, not
 * by the workspace. */
int cm_context_create(int device, void* stream, size_t workspace_bytes, cm_context** out);
/* Waits for the context's work and frees its memory and (if it made it) its
 * stream.  Sketches made on it stay valid. */
int cm_context_destroy(cm_context* ctx);
/* cm_compress_ex, cm_decompress and cm_recover on a context: the same
 * arguments and results (replacing the same reference functions), on the
 * context's device and stream.  Compress from host operands, decompress
 * into host memory and recover are synchronous; compress of device operands
 * and decompress into device memory are ordered on the stream
 * (cm_ctx_synchronize waits for them). */
int cm_ctx_compress(cm_context* ctx, int variant, int dtype, int n_a, int n_inner, int n_b,
                    const void* AT, int64_t ld_at, const void* B, int64_t ld_b, int d, int b,
                    uint64_t seed, int family, int mem, unsigned flags, cm_sketch** inout);
int cm_ctx_decompress(cm_context* ctx, const cm_sketch* sk, void* C, int64_t ldc, int mem);
int cm_ctx_recover(cm_context* ctx, const cm_sketch* sk, double tau, int64_t max_out,
                   int32_t* out_i, int32_t* out_j, double* out_val, int64_t* count);
int cm_ctx_synchronize(cm_context* ctx);

/* ---------------------------------------------------------------- memory
 * Device buffers for callers that keep inputs resident in HBM. */
int cm_dev_alloc(size_t bytes, void** ptr);
int cm_dev_free(void* ptr);
int cm_memcpy(void* dst, const void* src, size_t bytes, int dst_mem, int src_mem);
int cm_synchronize(void);

/* ---------------------------------------------------------------- timing
 * When enabled, the library brackets every kernel launch with HIP events on
 * the stream it runs on (its own stream, a batch's, or a server wave's);
 * cm_timing_get returns the summed device time (ms) and the launch count of
 * the kernel class `name` ("compress", "reduce", "combine", "decompress",
 * "hash", "plain_sketch", "plain_transform", "plain_mac"; batches and server
 * waves: "hash", "balance", "lists", "compress_fft", "compress_fwht",
 * "finish", "decompress"). */
int cm_timing_enable(int on);
/* Restricts the timing to the kernel class `name`, or to every class that
 * starts with its prefix when it ends in '*' (NULL or "" restores every
 * class).  Each event pair stalls the stream briefly, so a throughput
 * measurement times only its dominant class.  No reference counterpart: it
 * serves the measurement harness, as cm_timing_enable does. */
int cm_timing_only(const char* name);
int cm_timing_reset(void);
int cm_timing_get(const char* name, double* total_ms, int64_t* launches);

/* ------------------------------------------------------------ calibration
 * Kernels of known work that give the rates the device sustains, so that a
 * kernel's roofline fraction can be read against them as well as against
 * the published peaks.  No reference counterpart: they serve the
 * measurement harness.  Each runs one untimed launch, then `reps` launches
 * between two HIP events on the library stream; *ms is the mean per launch.
 *
 * cm_probe_alu: `blocks` workgroups of 256 threads, each thread `iters`
 * trips of 16 independent adds (x + c) or FMAs (fma(x, a, c)); *flops is the
 * work of one launch (an FMA counts 2).  Packed FP32 instructions are not
 * used.  blocks * iters is at most 2^26 and reps at most 100 (both probes),
 * so no call runs for long.
 *
 * cm_probe_stream: 16-byte accesses over `bytes` (a multiple of 16) of
 * 16-byte aligned device memory: copy src to dst (non-overlapping), read
 * src, or fill dst; *moved is the bytes read plus written by one launch.
 * Buffers that are not device allocations of at least `bytes` are rejected
 * with CM_EINVAL before any launch. */
#define CM_PROBE_F32_ADD 0
#define CM_PROBE_F32_FMA 1
#define CM_PROBE_F64_ADD 2
#define CM_PROBE_F64_FMA 3
#define CM_PROBE_COPY 0
#define CM_PROBE_READ 1
#define CM_PROBE_FILL 2
int cm_probe_alu(int kind, int blocks, int iters, int reps, double* ms, double* flops);
int cm_probe_stream(int kind, void* dst, const void* src, size_t bytes, int reps, double* ms,
                    double* moved);

/* ------------------------------------------------------------------ batches
 * Many independent FP32 products of uneven shape in one pass.  The reference
 * has no counterpart: it computes one product per call
 * (compmatmul.cpp:107-186, matmul_fft.hpp:155-178).  Each product's result
 * is the one cm_compress + cm_decompress would give for the same operands
 * (same hashes; the chunking of the inner dimension and so the FP32
 * summation order may differ).  Operands and outputs are device pointers.
 *
 * A plan owns every device buffer of the batch.  Running it launches one
 * schedule, hash and finishing launch each for all products, one persistent
 * compress launch per transform variant (workgroups drain XCD-local work
 * queues), and one decompression launch per (variant, d) class.  With
 * CM_BATCH_GRAPH the launch sequence is captured once into a HIP graph and
 * replayed. */
typedef struct cm_product {
  int variant;              /* CM_FFT or CM_FWHT */
  int m, n, p;              /* C is m x p; inner dimension n */
  const float* AT;          /* device, n x m, row stride ld_at >= m */
  int64_t ld_at;
  const float* B;           /* device, n x p, row stride ld_b >= p */
  int64_t ld_b;
  float* C;                 /* device, m x p, row stride ldc >= p; NULL: not written */
  int64_t ldc;
  int d, b;                 /* 1 <= d <= 16; b a power of two, 2 <= b <= 2^30 */
  uint64_t seed;            /* h1, h2, s1, s2 from seed + 0..3 */
  int family;               /* CM_MULTIPLY_ADD_SHIFT ... */
  double tau;               /* recover entries > tau; NaN: no recovery */
  int64_t max_out;          /* capacity of the product's recovered list */
} cm_product;

typedef struct cm_batch cm_batch;

#define CM_BATCH_GRAPH 1u /* capture the launches into a HIP graph on the first run */

/* Plan `count` products.  `grain` is the target work per compress work item
 * in k-steps of a product with m + p = 0 (the per-product inner-dimension
 * chunk is grain / (1 + (m + p) / 1700), at least 1); 0 = the autotuned value
 * for this batch's shapes (tuned on first use by timing the candidate
 * grains, then cached for the process). */
int cm_batch_plan(int count, const cm_product* products, unsigned flags, int grain,
                  cm_batch** out);
/* Run the plan (asynchronous on the library stream; cm_synchronize waits). */
int cm_batch_run(cm_batch* batch);
/* The grain the plan uses. */
int cm_batch_grain(const cm_batch* batch);
/* Recovered entries of product i, sorted by (row, column); *count is the
 * number found (may exceed max_out; only max_out are stored). */
int cm_batch_recovered(cm_batch* batch, int i, int64_t max_out, int32_t* out_i, int32_t* out_j,
                       float* out_val, int64_t* count);
/* Sketch values (d x b, FP32 on the host) of product i. */
int cm_batch_values(cm_batch* batch, int i, float* host_out);
void cm_batch_free(cm_batch* batch);

/* ------------------------------------------------------------------ serving
 * A request server on one GPU (no counterpart in the reference, which has
 * one blocking product per call).  A dispatcher thread admits queued
 * requests in arrival order while their device footprint fits the server's
 * HBM budget, and runs each admitted group as a batch ("wave") on a stream
 * of its own.  Up to max_waves waves are on the device at once, so requests
 * that arrive while others run are admitted as wave slots and memory free
 * up, and their work overlaps.  A request's hash arrays and sketch stay in
 * HBM until cm_server_release (cm_server_decompress reads them), and count
 * against the budget until then.  With a fixed grain a request's result
 * does not depend on which other requests share its wave. */
#define CM_DENSE 0
#define CM_CSR 1

typedef struct cm_request {
  int variant;                  /* CM_FFT or CM_FWHT */
  int m, n, p, d, b;
  uint64_t seed;
  int family;
  double tau;                   /* recover entries > tau (NaN: none) */
  int64_t max_out;              /* recovered entries kept */
  int format;                   /* CM_DENSE or CM_CSR, for both operands */
  /* CM_DENSE: FP32 row-major AT (n x m) and B (n x p) */
  const float* AT;
  const float* B;
  /* CM_CSR: AT and B as n-row CSR matrices (indptr[n + 1], indices, values) */
  const int32_t* at_indptr;
  const int32_t* at_indices;
  const float* at_values;
  const int32_t* b_indptr;
  const int32_t* b_indices;
  const float* b_values;
  float* C;                     /* m x p output, or NULL */
  int mem;                      /* where AT, B (or the CSR arrays) and C live:
                                   CM_HOST (0): uploaded by the server, C read
                                   back; CM_DEVICE: used in place (operands
                                   resident in HBM), C written directly.
                                   Device CSR arrays are not validated on the
                                   host; columns out of range are dropped. */
} cm_request;

typedef struct cm_server cm_server;

/* hbm_bytes: device memory the server may hold; max_waves: waves on the
 * device at once (1 .. 4, one stream each: HIP gives a process four hardware
 * queues by default); max_wave_products: requests per wave; grain as for
 * cm_batch_plan (must be > 0: results then depend on the request only). */
int cm_server_create(size_t hbm_bytes, int max_waves, int max_wave_products, int grain,
                     cm_server** out);
/* Queue a request (the descriptor is copied; the host buffers it points to
 * must stay valid until the request completes).  Returns its ticket. */
int cm_server_submit(cm_server* server, const cm_request* request, uint64_t* ticket);
/* Wait up to timeout_s seconds (< 0: forever): CM_OK when the request is
 * done, 1 on timeout, or the request's error status. */
int cm_server_wait(cm_server* server, uint64_t ticket, double timeout_s);
/* Recovered entries (sorted by row, column), the number found, and the
 * request's submit-to-done latency in seconds. */
int cm_server_result(cm_server* server, uint64_t ticket, int64_t max_out, int32_t* out_i,
                     int32_t* out_j, float* out_val, int64_t* count, double* latency_s);
/* Median-of-d estimate of all of C from the request's retained sketch. */
int cm_server_decompress(cm_server* server, uint64_t ticket, float* host_C);
/* The request's retained sketch values (d x b, FP32) copied to host memory.
 * Replaces: reading matmul_sketch<Hash>::p (matmul_sketch.hpp:80) of the
 * product the request describes. */
int cm_server_values(cm_server* server, uint64_t ticket, float* host_out);
/* Drop the request's retained sketch and its record. */
int cm_server_release(cm_server* server, uint64_t ticket);
typedef struct cm_server_stats {
  int64_t waves, requests_done, requests_failed;
  int max_waves_in_flight;
  size_t hbm_bytes, peak_bytes;
  /* dispatcher thread seconds: admitting, uploading, planning and launching
   * waves; reading back and retiring them */
  double launch_s, retire_s;
} cm_server_stats;
int cm_server_stats_get(cm_server* server, cm_server_stats* out);
/* Finish the waves in flight, fail the queued requests, free everything. */
int cm_server_destroy(cm_server* server);

/* Pinned host memory (fast, truly asynchronous uploads for server clients). */
int cm_host_alloc(size_t bytes, void** ptr);
int cm_host_free(void* ptr);

#ifdef __cplusplus
}
#endif
#endif /* COMPMATMUL_AMD_H */
