// Unfused "plain" GPU pipeline and the FP64 GEMM.
//
// The plain path runs the reference's stages as separate passes over HBM
// (matmul_fft.hpp:70-107, matmul_fwht.hpp:64-95): count-sketch every row k
// into HBM, transform every sketch in place, then reduce the pointwise
// products over k.  It materialises n x b sketches and spectra and is the
// per-product baseline the fused path is measured against, so every stage
// is written to stream HBM:
//   sketch     one workgroup per (row, range of at most 16 KB of buckets):
//              LDS scatter of the keys falling in the range, one coalesced
//              write of the range;
//   transform  in place per row, radix-16 (last pass narrower) register
//              passes over an LDS copy of the row; a row larger than 128 KB
//              runs as b = b1 * b2 in two passes, length-b1 columns (with
//              the twiddles W_b^(n2 k1)) then length-b2 rows;
//   MAC        tiles of (frequency range, chunk of k, repetition) summed in
//              registers, then the chunks summed in a fixed order into the
//              spectrum of each repetition (deterministic).
// It uses the same pair packing and balance (z = 2^e S a + i 2^-e S' b,
// accumulate z^2) and the same reduce/combine kernels as the fused path, so
// the two differ only in fusion.  Any b up to the square of the 128 KB tile
// length (2^28 for the FP32 FFT).
//
// cm::launch_gemm_f64 is the exact FP64 product behind compmatmul.gemm
// (array.hpp:81-95: C = AT^T B, all row-major).
#include <type_traits>

#include "device_common.hpp"
#include "kernels.hpp"
#include "lds_transform.hpp"

namespace cm {

namespace {

constexpr int NT = 1024;       // threads of the transform kernels
constexpr int ST = 256;        // threads of the sketch kernel
constexpr int MT = 256;        // threads of the MAC kernels

using namespace ldsx;

// log2 of the elements of one 128 KB transform tile (one workgroup per CU)
template <typename T, bool FWHT>
constexpr int lmax_log() {
  return sizeof(Elem<T, FWHT>) == 4 ? 15 : sizeof(Elem<T, FWHT>) == 8 ? 14 : 13;
}

// log2 of the buckets of one 16 KB sketch range
template <typename T, bool FWHT>
constexpr int range_log() {
  return sizeof(Elem<T, FWHT>) == 4 ? 12 : sizeof(Elem<T, FWHT>) == 8 ? 11 : 10;
}

template <typename T, bool FWHT>
__device__ __forceinline__ Elem<T, FWHT>* row_base(const PlainArgs& a, bool side_b, int64_t row) {
  return static_cast<Elem<T, FWHT>*>(side_b ? a.SB : a.SA) + (row << a.lb);
}

// Sketch of row (t, k) for the buckets [q R, (q + 1) R) of one workgroup:
// LDS scatter, then one coalesced write.  FWHT: blockIdx.y selects A (0) or
// B (1); FFT: the balanced pair z = 2^e S a + i 2^-e S' b.
template <typename T, bool FWHT>
__global__ void __launch_bounds__(ST) k_plain_sketch(PlainArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using E = Elem<T, FWHT>;
  E* z = reinterpret_cast<E*>(smem_raw);
  const int lr = min(a.lb, range_log<T, FWHT>());
  const int R = 1 << lr;
  const int q = blockIdx.x & ((1 << (a.lb - lr)) - 1);
  const int row = blockIdx.x >> (a.lb - lr);
  const int t = a.t0 + row / a.nk, k = a.k0 + row % a.nk;
  const bool side_b = FWHT && blockIdx.y != 0;
  constexpr int V = 16 / int(sizeof(E));
  using Vc = Vec16<E>;
  if (R >= V) {
    for (int i = threadIdx.x; i < R / V; i += ST) reinterpret_cast<Vc*>(z)[i] = Vc{};
  } else {
    for (int e = threadIdx.x; e < R; e += ST) z[e] = E{};
  }
  __syncthreads();
  const T* __restrict__ ar = static_cast<const T*>(a.AT) + int64_t(k) * a.ld_at;
  const T* __restrict__ br = static_cast<const T*>(a.B) + int64_t(k) * a.ld_b;
  const uint32_t* __restrict__ hA = a.hA + int64_t(t) * a.m;
  const uint32_t* __restrict__ hB = a.hB + int64_t(t) * a.p;
  const uint32_t qq = uint32_t(q);
  if constexpr (FWHT) {
    const int keys = side_b ? a.p : a.m;
    const T* __restrict__ v = side_b ? br : ar;
    const uint32_t* __restrict__ hh = side_b ? hB : hA;
    for (int i = threadIdx.x; i < keys; i += ST) {
      const uint32_t e = hh[i], h = hbucket(e);
      if ((h >> lr) == qq) atomicAdd(&z[h & (R - 1)], apply_sign(v[i], hneg(e)));
    }
  } else {
    const cpx<T> bal = static_cast<const cpx<T>*>(a.scale)[k];
    for (int i = threadIdx.x; i < a.m; i += ST) {
      const uint32_t e = hA[i], h = hbucket(e);
      if ((h >> lr) == qq) atomicAdd(&z[h & (R - 1)].x, apply_sign(ar[i] * bal.x, hneg(e)));
    }
    for (int j = threadIdx.x; j < a.p; j += ST) {
      const uint32_t e = hB[j], h = hbucket(e);
      if ((h >> lr) == qq) atomicAdd(&z[h & (R - 1)].y, apply_sign(br[j] * bal.y, hneg(e)));
    }
  }
  __syncthreads();
  E* out = row_base<T, FWHT>(a, side_b, row) + (int64_t(q) << lr);
  if (R >= V) {
    for (int i = threadIdx.x; i < R / V; i += ST) reinterpret_cast<Vc*>(out)[i] = reinterpret_cast<const Vc*>(z)[i];
  } else {
    for (int e = threadIdx.x; e < R; e += ST) out[e] = z[e];
  }
}

// Transform segments of length 2^l in place: segment `seg` of row `row` is
// elements [seg 2^l, (seg + 1) 2^l).  A whole row when b fits a tile; the
// second (row) pass of a two-pass transform otherwise.  Written back in
// natural order within the segment.
template <typename T, bool FWHT>
__global__ void __launch_bounds__(NT) k_plain_rows(PlainArgs a, int l) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using E = Elem<T, FWHT>;
  E* z = reinterpret_cast<E*>(smem_raw);
  const int segs_log = a.lb - l;
  const int64_t row = blockIdx.x >> segs_log;
  const int seg = blockIdx.x & ((1 << segs_log) - 1);
  E* g = row_base<T, FWHT>(a, FWHT && blockIdx.y != 0, row) + (int64_t(seg) << l);
  const int L = 1 << l;
  cpx<T>* lo = twiddle_lds<T, FWHT>(smem_raw, l);
  cpx<T>* hi = lo + TW_LO;
  if (!FWHT) fill_twiddles<NT, T>(lo, hi, l, static_cast<const cpx<T>*>(a.tw), a.lb);
  constexpr int V = 16 / int(sizeof(E));
  constexpr int U = 4;  // vector loads in flight per thread
  if (L >= V) {  // 16-byte vectors; the V elements of one share a padded run
    const int nv = L / V;
    const Vec16<E>* gv = reinterpret_cast<const Vec16<E>*>(g);
    for (int i0 = threadIdx.x; i0 < nv; i0 += NT * U) {
      Vec16<E> x[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + u * NT < nv) x[u] = gv[i0 + u * NT];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (i0 + u * NT < nv)
#pragma unroll
          for (int v = 0; v < V; ++v) z[pad((i0 + u * NT) * V + v)] = x[u].e[v];
    }
  } else {
    for (int e = threadIdx.x; e < L; e += NT) z[pad(e)] = g[e];
  }
  lds_transform<NT, T, FWHT>(z, l, 1, 0, lo, hi);
  if (L >= V) {
    const int nv = L / V;
    Vec16<E>* gv = reinterpret_cast<Vec16<E>*>(g);
    for (int i = threadIdx.x; i < nv; i += NT) {
      Vec16<E> x;
#pragma unroll
      for (int v = 0; v < V; ++v) x.e[v] = z[pad(FWHT ? i * V + v : dif_pos(i * V + v, l, rmax<T, FWHT>()))];
      gv[i] = x;
    }
  } else {
    for (int e = threadIdx.x; e < L; e += NT) g[e] = z[pad(FWHT ? e : dif_pos(e, l, rmax<T, FWHT>()))];
  }
}

// First pass of a two-pass transform, b = b1 b2 (b1 = 2^l1): length-b1
// transforms over n1 of the columns n2 in [c0, c0 + 2^lc) of one row,
// element n1 b2 + n2; the FFT output k1 is multiplied by W_b^(n2 k1) and
// stored at k1 b2 + n2.  The tile is column-major in LDS.
template <typename T, bool FWHT>
__global__ void __launch_bounds__(NT) k_plain_cols(PlainArgs a, int l1, int lc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  using E = Elem<T, FWHT>;
  E* z = reinterpret_cast<E*>(smem_raw);
  const int l2 = a.lb - l1;
  const int tiles_log = l2 - lc;
  const int64_t row = blockIdx.x >> tiles_log;
  const int c0 = (blockIdx.x & ((1 << tiles_log) - 1)) << lc;
  E* g = row_base<T, FWHT>(a, FWHT && blockIdx.y != 0, row);
  const int n = 1 << (l1 + lc), C = 1 << lc;
  constexpr int V = 16 / int(sizeof(E));
  constexpr int U = 4;  // vector loads in flight per thread
  if (C >= V) {  // 16-byte vectors of V adjacent columns
    const int nv = n / V;
    for (int i0 = threadIdx.x; i0 < nv; i0 += NT * U) {
      Vec16<E> x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = (i0 + u * NT) * V;
        if (i0 + u * NT < nv)
          x[u] = *reinterpret_cast<const Vec16<E>*>(g + (int64_t(e >> lc) << l2) + c0 + (e & (C - 1)));
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = (i0 + u * NT) * V;
        if (i0 + u * NT < nv)
#pragma unroll
          for (int v = 0; v < V; ++v) z[pad((((e & (C - 1)) + v) << l1) + (e >> lc))] = x[u].e[v];
      }
    }
  } else {
    for (int e = threadIdx.x; e < n; e += NT) {
      const int n1 = e >> lc, cc = e & (C - 1);
      z[pad((cc << l1) + n1)] = g[(int64_t(n1) << l2) + c0 + cc];
    }
  }
  const cpx<T>* tw = static_cast<const cpx<T>*>(a.tw);
  cpx<T>* lo = twiddle_lds<T, FWHT>(smem_raw, l1 + lc);
  cpx<T>* hi = lo + TW_LO;
  if (!FWHT) fill_twiddles<NT, T>(lo, hi, l1, tw, a.lb);
  lds_transform<NT, T, FWHT>(z, l1, C, 1 << l1, lo, hi);
  for (int e = threadIdx.x; e < n; e += NT) {
    const int k1 = e >> lc, cc = e & (C - 1);
    if constexpr (FWHT) {
      g[(int64_t(k1) << l2) + c0 + cc] = z[pad((cc << l1) + k1)];
    } else {
      const cpx<T> v = z[pad((cc << l1) + dif_pos(k1, l1, rmax<T, FWHT>()))];
      g[(int64_t(k1) << l2) + c0 + cc] = cmul(v, tw[int64_t(c0 + cc) * k1]);
    }
  }
}

// elements per MAC thread: one 16-byte load per operand row
template <typename T, bool FWHT>
constexpr int mac_v() { return 16 / int(sizeof(Elem<T, FWHT>)); }

// Partial spectral sums over one chunk of k: blockIdx.x a range of MT * V
// stored positions, blockIdx.y the chunk, blockIdx.z the repetition within
// the pass; part[(chunk * nt + tz) b + P].
template <typename T, bool FWHT>
__global__ void __launch_bounds__(MT) k_plain_mac(PlainArgs a, int kper) {
  using E = Elem<T, FWHT>;
  constexpr int V = mac_v<T, FWHT>();
  const int64_t b = int64_t(1) << a.lb;
  const int64_t P0 = (int64_t(blockIdx.x) * MT + threadIdx.x) * V;
  if (P0 >= b) return;
  const int tz = blockIdx.z;
  const int k0 = blockIdx.y * kper;
  const int k1 = min(a.nk, k0 + kper);
  const int nv = int(min<int64_t>(V, b - P0));
  E s[V];
#pragma unroll
  for (int u = 0; u < V; ++u) s[u] = E{};
  auto mac = [&](const E& x, const E& y, E& acc) {
    if constexpr (FWHT)
      acc += x * y;
    else
      acc = cadd(acc, csqr(x));
  };
  if (nv == V) {  // whole 16-byte vectors, the loads of KU rows in flight
    using Vc = Vec16<E>;
    constexpr int KU = 4;
    int k = k0;
    for (; k + KU <= k1; k += KU) {
      Vc x[KU], y[KU];
#pragma unroll
      for (int u = 0; u < KU; ++u) {
        const int64_t row = int64_t(tz) * a.nk + k + u;
        x[u] = *reinterpret_cast<const Vc*>(row_base<T, FWHT>(a, false, row) + P0);
        if constexpr (FWHT) y[u] = *reinterpret_cast<const Vc*>(row_base<T, FWHT>(a, true, row) + P0);
      }
#pragma unroll
      for (int u = 0; u < KU; ++u)
#pragma unroll
        for (int v = 0; v < V; ++v) mac(x[u].e[v], FWHT ? y[u].e[v] : x[u].e[v], s[v]);
    }
    for (; k < k1; ++k) {
      const int64_t row = int64_t(tz) * a.nk + k;
      const Vc x = *reinterpret_cast<const Vc*>(row_base<T, FWHT>(a, false, row) + P0);
      Vc y = x;
      if constexpr (FWHT) y = *reinterpret_cast<const Vc*>(row_base<T, FWHT>(a, true, row) + P0);
#pragma unroll
      for (int v = 0; v < V; ++v) mac(x.e[v], y.e[v], s[v]);
    }
  } else {  // b below one vector
    for (int k = k0; k < k1; ++k) {
      const int64_t row = int64_t(tz) * a.nk + k;
      const E* x = row_base<T, FWHT>(a, false, row) + P0;
      const E* y = FWHT ? row_base<T, FWHT>(a, true, row) + P0 : x;
#pragma unroll
      for (int u = 0; u < V; ++u)
        if (u < nv) mac(x[u], y[u], s[u]);
    }
  }
  E* out = static_cast<E*>(a.part) + (int64_t(blockIdx.y) * a.nt + tz) * b + P0;
#pragma unroll
  for (int u = 0; u < V; ++u)
    if (u < nv) out[u] = s[u];
}

// Chunks of one pass summed in order into the spectrum of each repetition,
// at its slot of the sliced layout the reduce kernel reads ([t][s][f] with
// X[s + b1 f] for the FFT, natural for the FWHT); added to what earlier
// passes over the same repetition left there unless a.first.  l1 > 0: the
// FFT rows came from the two-pass transform, so stored position P holds
// X[(P >> l2) + ((P mod b2) << l1)].
template <typename T, bool FWHT>
__global__ void __launch_bounds__(MT) k_plain_mac_sum(PlainArgs a, int kchunks, int l1) {
  using E = Elem<T, FWHT>;
  const int64_t b = int64_t(1) << a.lb;
  const int64_t P = int64_t(blockIdx.x) * MT + threadIdx.x;
  if (P >= b) return;
  const int tz = blockIdx.y;
  int64_t F = P;
  if (!FWHT && l1 > 0) {
    const int l2 = a.lb - l1;
    F = (P >> l2) + ((P & ((int64_t(1) << l2) - 1)) << l1);
  }
  int64_t slot = F;
  if (!FWHT) {
    const int lbs = a.lb - a.lb2;
    slot = ((F & ((int64_t(1) << lbs) - 1)) << a.lb2) + (F >> lbs);
  }
  E* acc = static_cast<E*>(a.acc) + int64_t(a.t0 + tz) * b + slot;
  const E* part = static_cast<const E*>(a.part) + int64_t(tz) * b + P;
  E s = a.first ? E{} : *acc;
  for (int c = 0; c < kchunks; ++c) {
    if constexpr (FWHT)
      s += part[int64_t(c) * a.nt * b];
    else
      s = cadd(s, part[int64_t(c) * a.nt * b]);
  }
  *acc = s;
}

template <typename K>
bool lds_ok(K kernel, size_t bytes) {
  return bytes <= 65536 ||
         hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                             int(bytes)) == hipSuccess;
}

template <typename T, bool FWHT>
hipError_t sketch_t(const PlainArgs& a, hipStream_t s) {
  const int lr = std::min(a.lb, range_log<T, FWHT>());
  const size_t lds = sizeof(Elem<T, FWHT>) << lr;
  const int64_t blocks = (int64_t(a.nt) * a.nk) << (a.lb - lr);
  k_plain_sketch<T, FWHT><<<dim3(unsigned(blocks), FWHT ? 2 : 1), dim3(ST), lds, s>>>(a);
  return hipGetLastError();
}

template <typename T, bool FWHT>
hipError_t transform_t(const PlainArgs& a, hipStream_t s) {
  constexpr int LM = lmax_log<T, FWHT>();
  // LDS sized to the segment (so rows of up to 64 KB keep two workgroups a CU)
  if (!lds_ok(&k_plain_rows<T, FWHT>, transform_bytes<T, FWHT>(LM)) ||
      !lds_ok(&k_plain_cols<T, FWHT>, transform_bytes<T, FWHT>(LM)))
    return hipErrorInvalidValue;
  const int64_t rows = int64_t(a.nt) * a.nk;
  const int l2 = std::min(a.lb, LM);
  const int l1 = a.lb - l2;
  if (l1 > 0) {
    const int lc = LM - l1;
    k_plain_cols<T, FWHT><<<dim3(unsigned(rows << (l2 - lc)), FWHT ? 2 : 1), dim3(NT),
                            transform_bytes<T, FWHT>(LM), s>>>(a, l1, lc);
  }
  k_plain_rows<T, FWHT><<<dim3(unsigned(rows << l1), FWHT ? 2 : 1), dim3(NT), transform_bytes<T, FWHT>(l2),
                          s>>>(a, l2);
  return hipGetLastError();
}

template <typename T, bool FWHT>
hipError_t mac_t(const PlainArgs& a, hipStream_t s) {
  constexpr int LM = lmax_log<T, FWHT>();
  constexpr int V = mac_v<T, FWHT>();
  const int64_t b = int64_t(1) << a.lb;
  const int kper = (a.nk + a.kchunks - 1) / a.kchunks;
  const int kchunks = (a.nk + kper - 1) / kper;
  const dim3 grid(unsigned((b + MT * V - 1) / (MT * V)), unsigned(kchunks), unsigned(a.nt));
  k_plain_mac<T, FWHT><<<grid, dim3(MT), 0, s>>>(a, kper);
  const int l1 = a.lb > LM ? a.lb - LM : 0;
  k_plain_mac_sum<T, FWHT><<<dim3(unsigned((b + MT - 1) / MT), unsigned(a.nt)), dim3(MT), 0, s>>>(a, kchunks, l1);
  return hipGetLastError();
}

// C[i][j] = sum_k AT[k][i] B[k][j] (FP64, all row-major).  A workgroup of
// 256 threads owns a GT x GT tile of C (GT = 128, or 64 where 128-tiles
// would leave CUs idle); thread (tx, ty) accumulates the (GT/16)^2 outputs
// at rows 2 ty + 32 r + u and columns 2 tx + 32 c + v (r, c < GT/32; u, v
// < 2) in registers.  Tiles of 8 inner indices of AT and B pass through two
// LDS buffers (k-major rows of GT), loaded one step ahead into registers;
// each thread reads its operands as 16-byte pairs, lanes 16 bytes apart
// (conflict-free).  Each output is one sequential FMA chain over k (the
// same order for any tiling), padded loads read zeros.
constexpr int GK = 8;

template <int GT>
__global__ void __launch_bounds__(256) k_gemm_f64(int m, int n, int p,
                                                  const double* __restrict__ AT,
                                                  const double* __restrict__ B,
                                                  double* __restrict__ C) {
  constexpr int NO = GT / 16;  // outputs per thread in each direction
  constexpr int H = GT / 64;   // pairs of each operand a thread moves per step
  __shared__ __align__(16) double sa[2][GK][GT];
  __shared__ __align__(16) double sb[2][GK][GT];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
  // global -> LDS: row kr of the step, columns 2 cq + 64 h
  const int kr = tid >> 5, cq = tid & 31;
  double2 ra[H], rb[H];
  auto load = [&](int k0) {
    const int k = k0 + kr;
#pragma unroll
    for (int h = 0; h < H; ++h) {
      const int ia = i0 + 2 * cq + 64 * h, jb = j0 + 2 * cq + 64 * h;
      double2 a = make_double2(0.0, 0.0), bv = make_double2(0.0, 0.0);
      if (k < n) {
        const double* arow = AT + int64_t(k) * m;
        const double* brow = B + int64_t(k) * p;
        if (ia + 1 < m && ((m & 1) == 0)) {
          a = *reinterpret_cast<const double2*>(arow + ia);
        } else {
          if (ia < m) a.x = arow[ia];
          if (ia + 1 < m) a.y = arow[ia + 1];
        }
        if (jb + 1 < p && ((p & 1) == 0)) {
          bv = *reinterpret_cast<const double2*>(brow + jb);
        } else {
          if (jb < p) bv.x = brow[jb];
          if (jb + 1 < p) bv.y = brow[jb + 1];
        }
      }
      ra[h] = a;
      rb[h] = bv;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int h = 0; h < H; ++h) {
      *reinterpret_cast<double2*>(&sa[buf][kr][2 * cq + 64 * h]) = ra[h];
      *reinterpret_cast<double2*>(&sb[buf][kr][2 * cq + 64 * h]) = rb[h];
    }
  };
  double acc[NO][NO];
#pragma unroll
  for (int r = 0; r < NO; ++r)
#pragma unroll
    for (int c = 0; c < NO; ++c) acc[r][c] = 0.0;
  const int steps = (n + GK - 1) / GK;
  if (steps > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int t = 0; t < steps; ++t) {
    const int buf = t & 1;
    if (t + 1 < steps) load((t + 1) * GK);
#pragma unroll
    for (int kk = 0; kk < GK; ++kk) {
      double a[NO], bv[NO];
#pragma unroll
      for (int q = 0; q < NO / 2; ++q) {
        const double2 x = *reinterpret_cast<const double2*>(&sa[buf][kk][2 * ty + 32 * q]);
        const double2 y = *reinterpret_cast<const double2*>(&sb[buf][kk][2 * tx + 32 * q]);
        a[2 * q] = x.x;
        a[2 * q + 1] = x.y;
        bv[2 * q] = y.x;
        bv[2 * q + 1] = y.y;
      }
#pragma unroll
      for (int r = 0; r < NO; ++r)
#pragma unroll
        for (int c = 0; c < NO; ++c) acc[r][c] = fma(a[r], bv[c], acc[r][c]);
    }
    if (t + 1 < steps) store(buf ^ 1);  // the other buffer was last read before the previous barrier
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < NO; ++r) {
    const int i = i0 + 2 * ty + 32 * (r >> 1) + (r & 1);
    if (i >= m) continue;
#pragma unroll
    for (int c = 0; c < NO; ++c) {
      const int j = j0 + 2 * tx + 32 * (c >> 1) + (c & 1);
      if (j < p) C[int64_t(i) * p + j] = acc[r][c];
    }
  }
}


}  // namespace

int plain_max_lb(int fwht, int f64) {
  const int lm = fwht ? (f64 ? lmax_log<double, true>() : lmax_log<float, true>())
                      : (f64 ? lmax_log<double, false>() : lmax_log<float, false>());
  return 2 * lm;
}

hipError_t launch_plain_sketch(const PlainArgs& a, hipStream_t s) {
  if (a.nk == 0 || a.nt == 0) return hipSuccess;
  if (a.f64) return a.fwht ? sketch_t<double, true>(a, s) : sketch_t<double, false>(a, s);
  return a.fwht ? sketch_t<float, true>(a, s) : sketch_t<float, false>(a, s);
}

hipError_t launch_plain_transform(const PlainArgs& a, hipStream_t s) {
  if (a.nk == 0 || a.nt == 0) return hipSuccess;
  if (a.f64) return a.fwht ? transform_t<double, true>(a, s) : transform_t<double, false>(a, s);
  return a.fwht ? transform_t<float, true>(a, s) : transform_t<float, false>(a, s);
}

hipError_t launch_plain_mac(const PlainArgs& a, hipStream_t s) {
  if (a.nk == 0 || a.nt == 0) return hipSuccess;
  if (a.f64) return a.fwht ? mac_t<double, true>(a, s) : mac_t<double, false>(a, s);
  return a.fwht ? mac_t<float, true>(a, s) : mac_t<float, false>(a, s);
}

hipError_t launch_gemm_f64(int m, int n, int p, const double* AT, const double* B, double* C,
                           hipStream_t s) {
  if (m == 0 || p == 0) return hipSuccess;
  // 128-tiles where they give every CU one, else 64-tiles
  const int64_t big = int64_t((p + 127) / 128) * ((m + 127) / 128);
  if (big >= 256) {
    k_gemm_f64<128><<<dim3(unsigned((p + 127) / 128), unsigned((m + 127) / 128)), dim3(256), 0, s>>>(m, n, p, AT, B, C);
  } else {
    k_gemm_f64<64><<<dim3(unsigned((p + 63) / 64), unsigned((m + 63) / 64)), dim3(256), 0, s>>>(m, n, p, AT, B, C);
  }
  return hipGetLastError();
}

}  // namespace cm
