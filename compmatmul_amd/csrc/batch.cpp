// Batch engine: many FP32 compressed products in one pass (cm_batch_*).
//
// Plan: validate every product, draw its hash parameters on the host, carve
// one device arena, and upload the tables that map workgroups and work items
// to products:
//   hash        a range of workgroups per product          (k_hash_batch)
//   balance     FFT products: a workgroup per four inner indices
//                                                         (k_pair_scale_batch)
//   schedules   one workgroup per (product, scatter list)  (k_*_lists_batch)
//   compress    (product, unit) work items in eight XCD queues per variant,
//               drained by one persistent launch per variant
//   finish      chunk sums, per-slice inverses, combines (one launch each)
//   decompress  one launch per (variant, d, form) class
// Run: that launch sequence on the library stream, captured once into a HIP
// graph with CM_BATCH_GRAPH.  Counters are cleared by a fill kernel, so the
// graph holds no memset node.
//
// Work items: product q is cut into nchunks_q chunks of kper_q inner indices;
// a unit is (chunk, repetition, slice) and costs about kper_q * c_q with
// c_q = 1 + (m_q + p_q) / 1700 (one k-step of both transforms plus the
// scatter of m + p keys, in units of a bare transform pair).  kper_q depends
// only on the product, the grain and the device (a product that cannot fill
// the GPU alone at that grain is cut into at most one round of resident
// workgroups), so a product's result does not depend on what else is in the
// batch.  Chunks go to the queue with the least work
// (largest units first), keeping all units of a chunk on one XCD so they
// share its L2 copy of the chunk's rows of AT and B.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/compmatmul_amd.h"
#include "batch_internal.hpp"
#include "recover_sort.hpp"
#include "host_hash.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

struct cm_batch {
  std::vector<cm_product> prods;
  unsigned flags = 0;
  int grain = 0;
  void* arena = nullptr;
  bool owns_arena = true;
  hipStream_t stream = nullptr;
  struct Prod {
    uint32_t* hA;
    uint32_t* hB;
    float* P;
    unsigned long long* counter;
    int32_t* oi;
    int32_t* oj;
    float* ov;
  };
  std::vector<Prod> pp;
  cm::HashArgs* hash = nullptr;
  int2* hash_blocks = nullptr;
  int n_hash_blocks = 0;
  cm::PairScaleArgs* pscale = nullptr;  // FFT products' balance scales (balance.hip)
  int2* pscale_blocks = nullptr;
  int n_pscale_blocks = 0;
  // per compress kernel kind (kind_of): FFT then FWHT, 1, 2 or 4 waves
  static constexpr int kKinds = 6;
  cm::RegJob* jobs[kKinds] = {};
  int2* list_blocks[kKinds] = {};
  int n_list_blocks[kKinds] = {};
  int2* items[kKinds] = {};
  int* qstart[kKinds] = {};
  int n_items[kKinds] = {};
  int grid[kKinds] = {};
  unsigned* counters = nullptr;  // 8 queue counters per kind, then a u64 per product
  int counter_words = 0;
  cm::FinishJob* finish = nullptr;
  int2* sum_blocks = nullptr;
  int2* red_blocks = nullptr;
  int2* comb_blocks = nullptr;
  int n_sum = 0, n_red = 0, n_comb = 0;
  size_t red_lds = 0;
  cm::DecompressArgs* dec = nullptr;
  struct DecClass {
    int fwht, d, form;
    size_t lds;
    int2* blocks;
    int nblocks;
  };
  std::vector<DecClass> classes;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

using cm::rt::error;

constexpr int kGrains[] = {16, 32, 64, 128, 256};
constexpr size_t kPartialCap = size_t(64) << 20;  // per product
std::map<uint64_t, int> g_tuned;                   // batch signature -> grain

int ilog2(int64_t b) {
  int l = 0;
  while ((int64_t(1) << l) < b) ++l;
  return l;
}

double step_cost(const cm_product& q) { return 1.0 + double(q.m + q.p) / 1700.0; }

constexpr int kKinds = cm_batch::kKinds;

// compress kernel of a product: 3 * fwht + log2(waves per workgroup), the
// slice-group kernels for sketches of at least two 2048-point slices
int kind_of(const cm_product& q) {
  const int lb = ilog2(q.b);
  const bool fw = q.variant == CM_FWHT;
  int logg = 0;
  if (fw ? cm::fwht_grp_supported(lb, q.m, q.p) : cm::fft_grp_supported(lb, q.m, q.p))
    logg = fw ? cm::fwht_grp_log_g(lb) : cm::fft_grp_log_g(lb);
  return (fw ? 3 : 0) + logg;
}

size_t kind_lists_bytes(int kind, const cm_product& q) {
  const int lb = ilog2(q.b);
  switch (kind) {
    case 0: return cm::fft_reg_lists_bytes(q.d, lb, q.m, q.p);
    case 3: return cm::fwht_reg_lists_bytes(q.d, lb, q.m, q.p);
    default:
      return kind < 3 ? cm::fft_grp_lists_bytes(q.d, lb, q.m, q.p) : cm::fwht_grp_lists_bytes(q.d, lb, q.m, q.p);
  }
}

int kind_slots(int kind) {
  const int logg = kind % 3, lb_of = logg == 1 ? 12 : 13;
  if (kind < 3) return logg ? cm::fft_grp_slots(lb_of) : cm::fft_reg_slots();
  return logg ? cm::fwht_grp_slots(lb_of) : cm::fwht_reg_slots();
}

uint64_t signature(int count, const cm_product* ps) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  mix(uint64_t(count));
  for (int i = 0; i < count; ++i) {
    const cm_product& q = ps[i];
    mix(uint64_t(q.variant));
    mix(uint64_t(q.m));
    mix(uint64_t(q.n));
    mix(uint64_t(q.p));
    mix(uint64_t(q.d));
    mix(uint64_t(q.b));
  }
  return h;
}

int validate(int i, const cm_product& q) {
  const std::string at = "product " + std::to_string(i) + ": ";
  if (q.variant != CM_FFT && q.variant != CM_FWHT)
    return error(CM_EINVAL, at + "variant must be CM_FFT or CM_FWHT");
  if (q.m < 0 || q.n < 0 || q.p < 0) return error(CM_EINVAL, at + "negative dimension");
  if (q.b < 2 || (q.b & (q.b - 1)) != 0)
    return error(CM_EINVAL, at + "b must be a power of two at least 2");
  if (q.d < 1 || q.d > 16) return error(CM_ENOTSUP, at + "batches take 1 <= d <= 16");
  if (q.family < CM_MULTIPLY_ADD_SHIFT || q.family > CM_TABULATION16)
    return error(CM_EHASH, at + "unknown hash family");
  if ((q.m > 0 && q.n > 0 && (!q.AT || q.ld_at < q.m)) || (q.p > 0 && q.n > 0 && (!q.B || q.ld_b < q.p)))
    return error(CM_EINVAL, at + "operand pointer or leading dimension");
  if (q.C && q.ldc < q.p) return error(CM_EINVAL, at + "ldc must be at least p");
  const int lb = ilog2(q.b);
  const bool ok = q.variant == CM_FWHT ? cm::fwht_reg_supported(lb, q.m, q.p)
                                       : cm::fft_reg_supported(lb, q.m, q.p);
  if (!ok) return error(CM_ENOTSUP, at + "shape outside the register kernels' range");
  if (!std::isnan(q.tau) && q.max_out < 0) return error(CM_EINVAL, at + "max_out must be >= 0");
  return CM_OK;
}

// device arena carved in two passes: sizes, then addresses
struct Carver {
  char* base = nullptr;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += std::max<size_t>(count, 1) * sizeof(T);
    return p;
  }
};

struct HostPlan {
  // per product
  std::vector<int> lbv, kper, nchunks, jobidx, kind;
  std::vector<std::vector<uint64_t>> mul;    // multiply-family parameters
  std::vector<std::vector<uint32_t>> tabs;   // tabulation table prefixes
  std::vector<int> tab_len;
};

int build_plan(int count, const cm_product* ps, unsigned flags, int grain, const cm::PlanEnv& env,
               cm_batch** out) {
  if (count < 0 || (count > 0 && !ps) || (!out && !env.size_only))
    return error(CM_EINVAL, "null or negative argument");
  const bool size_only = env.size_only != nullptr;
  for (int i = 0; i < count; ++i) {
    int rc = validate(i, ps[i]);
    if (rc) return rc;
  }
  int rc = cm::rt::init();
  if (rc) return rc;
  cm_batch* bt = new cm_batch();
  bt->prods.assign(ps, ps + count);
  bt->flags = flags;
  bt->grain = grain;
  bt->stream = env.stream ? env.stream : cm::rt::stream();
  HostPlan hp;
  hp.lbv.resize(count);
  hp.kper.resize(count);
  hp.nchunks.resize(count);
  hp.jobidx.resize(count);
  hp.kind.resize(count);
  hp.mul.resize(count);
  hp.tabs.resize(count);
  hp.tab_len.assign(count, 0);
  int njobs[kKinds] = {};
  int slots_of[kKinds] = {};
  for (int i = 0; i < count; ++i) {
    const cm_product& q = ps[i];
    const int fw = q.variant == CM_FWHT;
    const int lb = ilog2(q.b);
    hp.lbv[i] = std::max(lb, cm::kRegLb2);
    int kper = int(std::lround(double(grain) / step_cost(q)));
    kper = std::max(1, std::min(kper, std::max(1, q.n)));
    // A product too small to fill the GPU alone at grain-sized chunks is cut
    // finer, into at most one round of resident workgroups (the rule of the
    // single-product path): alone or in a small wave its units would
    // otherwise run long serial k loops on a few CUs.  The rule depends
    // only on the product and the device, not on the rest of the batch.
    if (q.n > 0) {
      const int kd = kind_of(q);
      if (!slots_of[kd]) slots_of[kd] = kind_slots(kd);
      const int64_t per_chunk = std::max<int64_t>(1, (int64_t(q.d) << cm::reg_lbs(lb)) >> (kd % 3));
      const int64_t fill = std::max<int64_t>(1, std::min<int64_t>(q.n, slots_of[kd] / per_chunk));
      kper = int(std::min<int64_t>(kper, (q.n + fill - 1) / fill));
    }
    const size_t chunk_bytes = (size_t(q.d) << hp.lbv[i]) * (fw ? 1 : 2) * sizeof(float);
    const int64_t max_chunks = std::max<int64_t>(1, int64_t(kPartialCap / chunk_bytes));
    if (q.n > 0 && (q.n + kper - 1) / kper > max_chunks) kper = int((q.n + max_chunks - 1) / max_chunks);
    hp.kper[i] = kper;
    hp.nchunks[i] = q.n > 0 ? (q.n + kper - 1) / kper : 0;
    hp.kind[i] = kind_of(q);
    hp.jobidx[i] = njobs[hp.kind[i]]++;
    if (size_only) {  // sizes of the parameter blocks only
      if (q.family == CM_MULTIPLY_ADD_SHIFT || q.family == CM_MULTIPLY_SHIFT) {
        hp.mul[i].resize(size_t(4) * q.d * (q.family == CM_MULTIPLY_ADD_SHIFT ? 2 : 1));
      } else {
        const int chars = q.family == CM_TABULATION8 ? 4 : 2;
        const int64_t maxkey = std::max(1, std::max(q.m, q.p)) - 1;
        const int64_t tl = q.family == CM_TABULATION8 ? 256 : std::min<int64_t>(65536, maxkey + 1);
        hp.tab_len[i] = int(tl);
        hp.tabs[i].resize(size_t(4) * q.d * chars * tl);
      }
    } else if (q.family == CM_MULTIPLY_ADD_SHIFT || q.family == CM_MULTIPLY_SHIFT) {
      std::vector<uint64_t> one;
      for (int f = 0; f < 4; ++f) {
        cm::draw_multiply(q.family, q.d, q.seed + uint64_t(f), one);
        hp.mul[i].insert(hp.mul[i].end(), one.begin(), one.end());
      }
    } else {
      const int chars = q.family == CM_TABULATION8 ? 4 : 2;
      const int64_t full = q.family == CM_TABULATION8 ? 256 : 65536;
      const int64_t maxkey = std::max(1, std::max(q.m, q.p)) - 1;
      const int64_t tl = q.family == CM_TABULATION8 ? 256 : std::min<int64_t>(full, maxkey + 1);
      hp.tab_len[i] = int(tl);
      std::vector<uint32_t> tabs;
      for (int f = 0; f < 4; ++f) {
        cm::draw_tables(q.family, q.d, q.seed + uint64_t(f), tabs);
        for (int t = 0; t < q.d; ++t)
          for (int c = 0; c < chars; ++c) {
            const uint32_t* src = tabs.data() + (int64_t(t) * chars + c) * full;
            hp.tabs[i].insert(hp.tabs[i].end(), src, src + tl);
          }
      }
    }
  }

  // host tables
  std::vector<cm::HashArgs> hash(count);
  std::vector<int2> hash_blocks;
  std::vector<cm::PairScaleArgs> pscale;
  std::vector<int2> pscale_blocks;
  std::vector<cm::RegJob> jobs[kKinds];
  for (int k = 0; k < kKinds; ++k) jobs[k].resize(njobs[k]);
  std::vector<int2> list_blocks[kKinds];
  std::vector<cm::FinishJob> fin(count);
  std::vector<int2> sum_blocks, red_blocks, comb_blocks;
  std::vector<cm::DecompressArgs> dec(count);

  const void* t2 = nullptr;
  std::vector<const void*> tw(count, nullptr);
  if (!size_only) {
    if (njobs[0] + njobs[1] + njobs[2] && (rc = cm::rt::fft_t2_table(&t2))) {
      delete bt;
      return rc;
    }
    for (int i = 0; i < count; ++i)
      if (ps[i].variant == CM_FFT && (rc = cm::rt::twiddles(hp.lbv[i], 0, &tw[i]))) {
        delete bt;
        return rc;
      }
  }

  // two passes over the arena: sizes (base == nullptr), then addresses
  Carver cv;
  std::vector<size_t> pinned_off;
  for (int pass = 0; pass < 2; ++pass) {
    cv.off = 0;
    bt->pp.assign(count, {});
    std::vector<void*> lists(count), partial(count), V(count), scale(count);
    std::vector<const uint64_t*> mulp(count);
    std::vector<const uint32_t*> tabp(count);
    for (int i = 0; i < count; ++i) {
      const cm_product& q = ps[i];
      const int fw = q.variant == CM_FWHT;
      auto& P = bt->pp[i];
      if (env.keep) {
        P.hA = env.keep[i].hA;
        P.hB = env.keep[i].hB;
        P.P = env.keep[i].P;
      } else {
        P.hA = cv.take<uint32_t>(size_t(q.d) * q.m);
        P.hB = cv.take<uint32_t>(size_t(q.d) * q.p);
        P.P = cv.take<float>(size_t(q.d) * q.b);
      }
      const int64_t mo = std::isnan(q.tau) ? 0 : q.max_out;
      P.oi = cv.take<int32_t>(size_t(mo));
      P.oj = cv.take<int32_t>(size_t(mo));
      P.ov = cv.take<float>(size_t(mo));
      lists[i] = cv.take<char>(kind_lists_bytes(hp.kind[i], q));
      const size_t chunk_floats = (size_t(q.d) << hp.lbv[i]) * (fw ? 1 : 2);
      partial[i] = cv.take<float>(size_t(std::max(1, hp.nchunks[i])) * chunk_floats);
      V[i] = cv.take<float>(chunk_floats);
      scale[i] = fw ? nullptr : cv.take<float2>(size_t(q.n));
      if (!hp.mul[i].empty() && hp.mul[i].size() > 256) mulp[i] = cv.take<uint64_t>(hp.mul[i].size());
      if (!hp.tabs[i].empty()) tabp[i] = cv.take<uint32_t>(hp.tabs[i].size());
    }
    bt->counter_words = 8 * kKinds + 2 * count;
    bt->counters = cv.take<unsigned>(size_t(bt->counter_words));
    // tables (filled below on the second pass, uploaded from one host image)
    const size_t tables_begin = (cv.off + 255) & ~size_t(255);
    bt->hash = cv.take<cm::HashArgs>(size_t(count));
    bt->pscale = cv.take<cm::PairScaleArgs>(size_t(count));
    bt->finish = cv.take<cm::FinishJob>(size_t(count));
    bt->dec = cv.take<cm::DecompressArgs>(size_t(count));
    for (int k = 0; k < kKinds; ++k) bt->jobs[k] = cv.take<cm::RegJob>(size_t(njobs[k]));

    // per-product tables and block maps
    hash_blocks.clear();
    pscale.clear();
    pscale_blocks.clear();
    sum_blocks.clear();
    red_blocks.clear();
    comb_blocks.clear();
    for (int k = 0; k < kKinds; ++k) list_blocks[k].clear();
    for (int i = 0; i < count; ++i) {
      const cm_product& q = ps[i];
      const int fw = q.variant == CM_FWHT;
      const int lb = ilog2(q.b);
      auto& P = bt->pp[i];
      cm::HashArgs& ha = hash[i];
      ha = cm::HashArgs{};
      ha.family = q.family;
      ha.d = q.d;
      ha.lb = lb;
      ha.m = q.m;
      ha.p = q.p;
      ha.outA = P.hA;
      ha.outB = P.hB;
      if (!hp.mul[i].empty()) {
        if (hp.mul[i].size() <= 256) {
          ha.use_inline = 1;
          std::copy(hp.mul[i].begin(), hp.mul[i].end(), ha.inline_params);
        } else {
          ha.mul_params = mulp[i];
        }
      } else {
        ha.tables = tabp[i];
        ha.tab_len = hp.tab_len[i];
      }
      const int hb = cm::hash_batch_blocks(ha);
      for (int g = 0; g < hb; ++g) hash_blocks.push_back(make_int2(i, g));

      cm::RegJob& j = jobs[hp.kind[i]][hp.jobidx[i]];
      j = cm::RegJob{};
      j.AT = q.AT;
      j.ld_at = q.ld_at;
      j.B = q.B;
      j.ld_b = q.ld_b;
      j.hA = P.hA;
      j.hB = P.hB;
      j.m = q.m;
      j.n = q.n;
      j.p = q.p;
      j.d = q.d;
      j.lb = lb;
      j.lbs = cm::reg_lbs(lb);
      j.kper = hp.kper[i];
      j.nchunks = hp.nchunks[i];
      j.lists = lists[i];
      j.tw = tw[i];
      j.t2 = t2;
      j.scale = static_cast<const float2*>(scale[i]);
      j.partial = static_cast<float*>(partial[i]);
      if (!fw && q.n > 0) {
        const cm::PairScaleArgs pa{0, q.AT, q.ld_at, q.B, q.ld_b, q.m, q.n, q.p, scale[i]};
        const int nb = cm::pair_scale_blocks(pa);
        for (int g = 0; g < nb; ++g) pscale_blocks.push_back(make_int2(int(pscale.size()), g));
        pscale.push_back(pa);
      }
      const int nl = cm::reg_lists_count(fw, q.d, lb);
      for (int g = 0; g < nl; ++g) list_blocks[hp.kind[i]].push_back(make_int2(hp.jobidx[i], g));

      cm::FinishJob& f = fin[i];
      const int f64r = q.n > 16384;
      f.r = cm::ReduceArgs{fw, 0, f64r, partial[i], hp.nchunks[i], q.d, hp.lbv[i], cm::kRegLb2, tw[i], V[i]};
      f.c = cm::CombineArgs{fw, 0, V[i], q.d, hp.lbv[i], cm::kRegLb2, tw[i], P.P, lb};
      int nsum = 0, nred = 0;
      const int ncomb = cm::finish_batch_blocks(f, &nsum, &nred);
      for (int g = 0; g < nsum; ++g) sum_blocks.push_back(make_int2(i, g));
      for (int g = 0; g < nred; ++g) red_blocks.push_back(make_int2(i, g));
      for (int g = 0; g < ncomb; ++g) comb_blocks.push_back(make_int2(i, g));

      cm::DecompressArgs& da = dec[i];
      da = cm::DecompressArgs{};
      da.fwht = fw;
      da.f64 = 0;
      da.P = P.P;
      da.hA = P.hA;
      da.hB = P.hB;
      da.m = q.m;
      da.p = q.p;
      da.d = q.d;
      da.b = q.b;
      da.C = q.C;
      da.ldc = q.ldc;
      P.counter = reinterpret_cast<unsigned long long*>(bt->counters + 8 * kKinds + 2 * i);
      if (!std::isnan(q.tau)) {
        da.recover = 1;
        da.tau = q.tau;
        da.counter = P.counter;
        da.max_out = q.max_out;
        da.oi = P.oi;
        da.oj = P.oj;
        da.ov = P.ov;
      }
    }
    bt->hash_blocks = cv.take<int2>(hash_blocks.size());
    bt->n_hash_blocks = int(hash_blocks.size());
    bt->pscale_blocks = cv.take<int2>(pscale_blocks.size());
    bt->n_pscale_blocks = int(pscale_blocks.size());
    for (int k = 0; k < kKinds; ++k) {
      bt->list_blocks[k] = cv.take<int2>(list_blocks[k].size());
      bt->n_list_blocks[k] = int(list_blocks[k].size());
    }
    bt->sum_blocks = cv.take<int2>(sum_blocks.size());
    bt->red_blocks = cv.take<int2>(red_blocks.size());
    bt->comb_blocks = cv.take<int2>(comb_blocks.size());
    bt->n_sum = int(sum_blocks.size());
    bt->n_red = int(red_blocks.size());
    bt->n_comb = int(comb_blocks.size());
    bt->red_lds = 0;
    for (int i = 0; i < count; ++i)
      bt->red_lds = std::max(bt->red_lds, cm::reduce_lds_bytes(ps[i].variant == CM_FWHT, sizeof(float), cm::kRegLb2));

    // compress work items: chunks to the least-loaded of eight queues
    std::vector<int2> items_host[kKinds];
    std::vector<int> qstart_host[kKinds];
    for (int kd = 0; kd < kKinds; ++kd) {
      struct Chunk {
        int job, chunk, units;
        double unit_cost;
      };
      std::vector<Chunk> chunks;
      for (int i = 0; i < count; ++i) {
        if (hp.kind[i] != kd) continue;
        // workgroups per chunk of k: one per (t, slice), or per group of
        // 2^(kd % 3) slices
        const int per_chunk = (ps[i].d << cm::reg_lbs(ilog2(ps[i].b))) >> (kd % 3);
        for (int c = 0; c < hp.nchunks[i]; ++c) {
          const int k = std::min(hp.kper[i], ps[i].n - c * hp.kper[i]);
          chunks.push_back({hp.jobidx[i], c, per_chunk, k * step_cost(ps[i])});
        }
      }
      std::stable_sort(chunks.begin(), chunks.end(),
                       [](const Chunk& a, const Chunk& b) { return a.unit_cost > b.unit_cost; });
      std::vector<std::vector<int2>> q(8);
      std::vector<double> load(8, 0.0);
      for (const Chunk& c : chunks) {
        const int x = int(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += c.units * c.unit_cost;
        for (int u = 0; u < c.units; ++u) q[x].push_back(make_int2(c.job, c.chunk * c.units + u));
      }
      qstart_host[kd].assign(9, 0);
      for (int x = 0; x < 8; ++x) {
        qstart_host[kd][x + 1] = qstart_host[kd][x] + int(q[x].size());
        items_host[kd].insert(items_host[kd].end(), q[x].begin(), q[x].end());
      }
      bt->items[kd] = cv.take<int2>(items_host[kd].size());
      bt->qstart[kd] = cv.take<int>(9);
      bt->n_items[kd] = int(items_host[kd].size());
      const int slots = items_host[kd].empty() ? 8 : kind_slots(kd);
      bt->grid[kd] = items_host[kd].empty() ? 0 : std::max(8, std::min(slots, (bt->n_items[kd] + 7) / 8 * 8) / 8 * 8);
    }

    // decompress classes
    std::map<std::tuple<int, int, int>, std::vector<int2>> cls;
    std::map<std::tuple<int, int, int>, size_t> cls_lds;
    for (int i = 0; i < count; ++i) {
      const cm_product& q = ps[i];
      const int fw = int(q.variant == CM_FWHT);
      const int form = cm::decompress_batch_form(fw, q.d, q.b);
      const auto key = std::make_tuple(fw, q.d, form);
      const int nb = cm::decompress_batch_blocks(q.d, form, q.m, q.p);
      auto& v = cls[key];
      for (int g = 0; g < nb; ++g) v.push_back(make_int2(i, g));
      cls_lds[key] = std::max(cls_lds[key], cm::decompress_batch_lds(form, q.b));
    }
    bt->classes.clear();
    std::vector<std::vector<int2>*> cls_host;
    for (auto& kv : cls) {
      cm_batch::DecClass c{std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
                           cls_lds[kv.first], cv.take<int2>(kv.second.size()), int(kv.second.size())};
      bt->classes.push_back(c);
      cls_host.push_back(&kv.second);
    }
    const size_t tables_end = cv.off;

    if (pass == 0) {
      if (size_only) {
        *env.size_only = std::max<size_t>(cv.off, 256);
        delete bt;
        return CM_OK;
      }
      if (env.arena) {
        if (env.arena_bytes < cv.off) {
          delete bt;
          return error(CM_ENOMEM, "batch arena too small: " + std::to_string(cv.off) + " > " +
                                      std::to_string(env.arena_bytes) + " bytes");
        }
        bt->arena = env.arena;
        bt->owns_arena = false;
      } else {
        const hipError_t e = hipMalloc(&bt->arena, std::max<size_t>(cv.off, 256));
        if (e != hipSuccess) {
          delete bt;
          return error(e == hipErrorOutOfMemory ? CM_ENOMEM : CM_EDEVICE,
                       std::string("batch arena of ") + std::to_string(cv.off) + " bytes: " +
                           hipGetErrorString(e));
        }
      }
      cv.base = static_cast<char*>(bt->arena);
      continue;
    }

    // second pass: one host image of [tables_begin, tables_end) plus the
    // hash parameter blocks, uploaded synchronously
    std::vector<char> img(tables_end - tables_begin, 0);
    auto put = [&](const void* dev, const void* src, size_t bytes) {
      if (!bytes) return;
      std::memcpy(img.data() + (static_cast<const char*>(dev) - cv.base - tables_begin), src, bytes);
    };
    put(bt->hash, hash.data(), hash.size() * sizeof(cm::HashArgs));
    put(bt->pscale, pscale.data(), pscale.size() * sizeof(cm::PairScaleArgs));
    put(bt->pscale_blocks, pscale_blocks.data(), pscale_blocks.size() * sizeof(int2));
    put(bt->finish, fin.data(), fin.size() * sizeof(cm::FinishJob));
    put(bt->dec, dec.data(), dec.size() * sizeof(cm::DecompressArgs));
    for (int k = 0; k < kKinds; ++k) {
      put(bt->jobs[k], jobs[k].data(), jobs[k].size() * sizeof(cm::RegJob));
      put(bt->list_blocks[k], list_blocks[k].data(), list_blocks[k].size() * sizeof(int2));
      put(bt->items[k], items_host[k].data(), items_host[k].size() * sizeof(int2));
      put(bt->qstart[k], qstart_host[k].data(), 9 * sizeof(int));
    }
    put(bt->hash_blocks, hash_blocks.data(), hash_blocks.size() * sizeof(int2));
    put(bt->sum_blocks, sum_blocks.data(), sum_blocks.size() * sizeof(int2));
    put(bt->red_blocks, red_blocks.data(), red_blocks.size() * sizeof(int2));
    put(bt->comb_blocks, comb_blocks.data(), comb_blocks.size() * sizeof(int2));
    for (size_t c = 0; c < bt->classes.size(); ++c)
      put(bt->classes[c].blocks, cls_host[c]->data(), cls_host[c]->size() * sizeof(int2));
    // stream-ordered uploads: from the caller's pinned staging when they
    // fit (nothing to wait for; the caller keeps the staging until the
    // stream has passed them), else from the pageable host images, waited
    // for before they go out of scope
    size_t upload = img.size();
    for (int i = 0; i < count; ++i) {
      if (hp.mul[i].size() > 256) upload += hp.mul[i].size() * 8;
      upload += hp.tabs[i].size() * 4;
    }
    const bool staged = env.staging && upload <= env.staging_bytes;
    size_t soff = 0;
    auto copy_in = [&](void* dst, const void* src, size_t bytes) {
      if (!staged) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, bt->stream);
      std::memcpy(env.staging + soff, src, bytes);
      const hipError_t r = hipMemcpyAsync(dst, env.staging + soff, bytes, hipMemcpyHostToDevice, bt->stream);
      soff += bytes;
      return r;
    };
    hipError_t e = copy_in(cv.base + tables_begin, img.data(), img.size());
    for (int i = 0; i < count && e == hipSuccess; ++i) {
      if (hp.mul[i].size() > 256)
        e = copy_in(const_cast<uint64_t*>(mulp[i]), hp.mul[i].data(), hp.mul[i].size() * 8);
      if (e == hipSuccess && !hp.tabs[i].empty())
        e = copy_in(const_cast<uint32_t*>(tabp[i]), hp.tabs[i].data(), hp.tabs[i].size() * 4);
    }
    if (e == hipSuccess && !staged) e = hipStreamSynchronize(bt->stream);
    if (e != hipSuccess) {
      if (bt->owns_arena) (void)hipFree(bt->arena);
      delete bt;
      return error(CM_EDEVICE, std::string("batch table upload: ") + hipGetErrorString(e));
    }
  }
  *out = bt;
  return CM_OK;
}

// the launch sequence of one run; `untimed` skips the per-kernel timing
// events (graph capture, autotuning)
int enqueue(cm_batch* bt, bool untimed) {
  hipStream_t s = bt->stream;
  auto run = [&](const char* name, const std::function<hipError_t()>& f) -> int {
    if (untimed) {
      const hipError_t e = f();
      return e == hipSuccess ? CM_OK
                             : error(CM_EDEVICE, std::string(name) + " launch failed: " + hipGetErrorString(e));
    }
    return cm::rt::timed_on(bt->stream, name, f);
  };
  int rc;
  if ((rc = run("hash", [&] { return cm::launch_hash_batch(bt->hash, bt->hash_blocks, bt->n_hash_blocks, s); })))
    return rc;
  if ((rc = run("zero", [&] { return cm::launch_zero_u32(bt->counters, bt->counter_words, s); })))
    return rc;
  if (bt->n_pscale_blocks &&
      (rc = run("balance", [&] {
         return cm::launch_pair_scale_batch(bt->pscale, bt->pscale_blocks, bt->n_pscale_blocks, s);
       })))
    return rc;
  for (int k = 0; k < kKinds; ++k) {
    if (!bt->n_items[k] && !bt->n_list_blocks[k]) continue;
    const int logg = k % 3;
    const bool fw = k >= 3;
    if ((rc = run("lists", [&] {
           if (logg == 0)
             return fw ? cm::launch_fwht_reg_lists_batch(bt->jobs[k], bt->list_blocks[k], bt->n_list_blocks[k], s)
                       : cm::launch_fft_reg_lists_batch(bt->jobs[k], bt->list_blocks[k], bt->n_list_blocks[k], s);
           return fw ? cm::launch_fwht_grp_lists_batch(logg, bt->jobs[k], bt->list_blocks[k], bt->n_list_blocks[k], s)
                     : cm::launch_fft_grp_lists_batch(logg, bt->jobs[k], bt->list_blocks[k], bt->n_list_blocks[k], s);
         })))
      return rc;
  }
  for (int k = 0; k < kKinds; ++k) {
    if (!bt->n_items[k]) continue;
    const int logg = k % 3;
    const bool fw = k >= 3;
    unsigned* ctr = bt->counters + 8 * k;
    if ((rc = run(fw ? "compress_fwht" : "compress_fft", [&] {
           if (logg == 0)
             return fw ? cm::launch_fwht_reg_batch(bt->jobs[k], bt->items[k], bt->qstart[k], ctr, bt->grid[k], s)
                       : cm::launch_fft_reg_batch(bt->jobs[k], bt->items[k], bt->qstart[k], ctr, bt->grid[k], s);
           return fw ? cm::launch_fwht_grp_batch(logg, bt->jobs[k], bt->items[k], bt->qstart[k], ctr, bt->grid[k], s)
                     : cm::launch_fft_grp_batch(logg, bt->jobs[k], bt->items[k], bt->qstart[k], ctr, bt->grid[k], s);
         })))
      return rc;
  }
  if ((rc = run("finish", [&] {
         return cm::launch_finish_batch(bt->finish, bt->sum_blocks, bt->n_sum, bt->red_blocks, bt->n_red,
                                        bt->red_lds, bt->comb_blocks, bt->n_comb, s);
       })))
    return rc;
  for (const auto& c : bt->classes)
    if ((rc = run("decompress", [&] {
           return cm::launch_decompress_batch(c.fwht, c.d, c.form, c.lds, bt->dec, c.blocks, c.nblocks, s);
         })))
      return rc;
  return CM_OK;
}

int run_locked(cm_batch* bt) {
  if (!(bt->flags & CM_BATCH_GRAPH)) return enqueue(bt, false);
  hipStream_t s = bt->stream;
  if (!bt->exec) {
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return error(CM_EDEVICE, std::string("begin capture: ") + hipGetErrorString(e));
    const int rc = enqueue(bt, true);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(s, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return error(CM_EDEVICE, std::string("end capture: ") + hipGetErrorString(e));
    bt->graph = graph;
    e = hipGraphInstantiate(&bt->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return error(CM_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  const hipError_t e = hipGraphLaunch(bt->exec, s);
  return e == hipSuccess ? CM_OK : error(CM_EDEVICE, std::string("graph launch: ") + hipGetErrorString(e));
}

void free_locked(cm_batch* bt) {
  if (!bt) return;
  (void)hipStreamSynchronize(bt->stream);
  if (bt->exec) (void)hipGraphExecDestroy(bt->exec);
  if (bt->graph) (void)hipGraphDestroy(bt->graph);
  if (bt->arena && bt->owns_arena) (void)hipFree(bt->arena);
  delete bt;
}

// Autotuner: time each candidate grain on this batch (one warm-up run, then
// the best of two timed runs) and keep the fastest.
int tune(int count, const cm_product* ps, unsigned flags, int* best) {
  hipStream_t s = cm::rt::stream();
  double best_ms = 1e300;
  *best = kGrains[0];
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return error(CM_EDEVICE, "event creation failed");
  int rc = CM_OK;
  for (int grain : kGrains) {
    cm_batch* bt = nullptr;
    if ((rc = build_plan(count, ps, flags & ~CM_BATCH_GRAPH, grain, cm::PlanEnv{}, &bt))) break;
    double ms_min = 1e300;
    for (int r = 0; r < 3 && rc == CM_OK; ++r) {
      (void)hipEventRecord(e0, s);
      rc = enqueue(bt, true);
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) rc = error(CM_EDEVICE, "autotune run failed");
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0) ms_min = std::min(ms_min, double(ms));
    }
    free_locked(bt);
    if (rc) break;
    if (ms_min < best_ms) {
      best_ms = ms_min;
      *best = grain;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace

namespace cm {

int batch_plan_env(int count, const cm_product* ps, unsigned flags, int grain, const PlanEnv& env,
                   cm_batch** out) {
  if (grain <= 0) return error(CM_EINVAL, "grain must be positive");
  return build_plan(count, ps, flags & ~CM_BATCH_GRAPH, grain, env, out);
}

int batch_enqueue(cm_batch* bt) { return enqueue(bt, false); }

ProductView batch_product(const cm_batch* bt, int i) {
  const auto& P = bt->pp[i];
  return ProductView{P.counter, P.oi, P.oj, P.ov};
}

const unsigned long long* batch_counters(const cm_batch* bt) {
  return reinterpret_cast<const unsigned long long*>(bt->counters + 8 * kKinds);
}

void batch_destroy(cm_batch* bt) {
  if (!bt) return;
  if (bt->arena && bt->owns_arena) (void)hipFree(bt->arena);
  delete bt;
}

}  // namespace cm

extern "C" {

int cm_batch_plan(int count, const cm_product* products, unsigned flags, int grain, cm_batch** out) {
  std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
  if (!out) return error(CM_EINVAL, "null output handle");
  if (grain < 0) return error(CM_EINVAL, "grain must be >= 0");
  *out = nullptr;
  if (grain == 0) {
    int rc = cm::rt::init();
    if (rc) return rc;
    const uint64_t sig = signature(count, products);
    auto it = g_tuned.find(sig);
    if (it != g_tuned.end()) {
      grain = it->second;
    } else {
      for (int i = 0; i < count && products; ++i)
        if ((rc = validate(i, products[i]))) return rc;
      if ((rc = tune(count, products, flags, &grain))) return rc;
      g_tuned[sig] = grain;
    }
  }
  return build_plan(count, products, flags, grain, cm::PlanEnv{}, out);
}

int cm_batch_run(cm_batch* batch) {
  if (!batch) return error(CM_EINVAL, "null batch");
  std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
  return run_locked(batch);
}

int cm_batch_grain(const cm_batch* batch) { return batch ? batch->grain : 0; }

int cm_batch_recovered(cm_batch* batch, int i, int64_t max_out, int32_t* out_i, int32_t* out_j,
                       float* out_val, int64_t* count) {
  if (!batch || !count) return error(CM_EINVAL, "null argument");
  if (i < 0 || i >= int(batch->prods.size())) return error(CM_EINVAL, "product index out of range");
  if (max_out < 0) return error(CM_EINVAL, "max_out must be non-negative");
  std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
  const cm_product& q = batch->prods[i];
  if (std::isnan(q.tau)) return error(CM_EINVAL, "the product was planned without recovery");
  const auto& P = batch->pp[i];
  hipStream_t s = batch->stream;
  unsigned long long found = 0;
  if (hipMemcpyAsync(&found, P.counter, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return error(CM_EDEVICE, "reading the recovery counter failed");
  *count = int64_t(found);
  const int64_t k = std::min<int64_t>({int64_t(found), max_out, q.max_out});
  if (k == 0) return CM_OK;
  std::vector<int32_t> vi(k), vj(k);
  std::vector<float> vv(k);
  if (hipMemcpy(vi.data(), P.oi, k * 4, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(vj.data(), P.oj, k * 4, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(vv.data(), P.ov, k * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return error(CM_EDEVICE, "reading the recovered list failed");
  cm::sort_recovered(k, q.p, vi.data(), vj.data(), vv.data());
  for (int64_t t = 0; t < k; ++t) {
    if (out_i) out_i[t] = vi[t];
    if (out_j) out_j[t] = vj[t];
    if (out_val) out_val[t] = vv[t];
  }
  return CM_OK;
}

int cm_batch_values(cm_batch* batch, int i, float* host_out) {
  if (!batch || !host_out) return error(CM_EINVAL, "null argument");
  if (i < 0 || i >= int(batch->prods.size())) return error(CM_EINVAL, "product index out of range");
  std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
  const cm_product& q = batch->prods[i];
  hipStream_t s = batch->stream;
  if (hipStreamSynchronize(s) != hipSuccess ||
      hipMemcpy(host_out, batch->pp[i].P, size_t(q.d) * q.b * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return error(CM_EDEVICE, "reading the sketch failed");
  return CM_OK;
}

void cm_batch_free(cm_batch* batch) {
  std::lock_guard<std::recursive_mutex> lk(cm::rt::mutex());
  free_locked(batch);
}

}  // extern "C"
