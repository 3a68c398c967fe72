// Ordering of recovered entries by (i, j) on the host.
//
// Recovery appends (i, j, value) through a device counter, in no particular
// order; callers receive the list sorted by (i, j), the row-major order of C.
// The key i * p + j (below m * p) is sorted with a least-significant-digit
// radix sort, 8 bits per pass over the key's significant bytes only, which
// carries the entries along: linear in the list length, where a comparison
// sort through an index array cost about 0.2 ms per served request with
// thousands of entries.  Keys are distinct, so the order is the same as any
// sort by (i, j).
#pragma once
#include <stdint.h>

#include <cstring>
#include <vector>

namespace cm {

template <typename V>
void sort_recovered(int64_t n, int64_t p, int32_t* oi, int32_t* oj, V* ov) {
  if (n <= 1) return;
  struct E {
    uint64_t key;
    V v;
  };
  const size_t len = static_cast<size_t>(n);
  std::vector<E> a(len), b(len);
  uint64_t top = 0;
  for (int64_t t = 0; t < n; ++t) {
    a[t].key = uint64_t(oi[t]) * uint64_t(p) + uint64_t(oj[t]);
    a[t].v = ov ? ov[t] : V(0);
    top |= a[t].key;
  }
  for (int shift = 0; shift < 64 && (top >> shift) != 0; shift += 8) {
    size_t count[257] = {0};
    for (const E& e : a) ++count[((e.key >> shift) & 0xff) + 1];
    for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
    for (const E& e : a) b[count[(e.key >> shift) & 0xff]++] = e;
    a.swap(b);
  }
  for (int64_t t = 0; t < n; ++t) {
    oi[t] = int32_t(a[t].key / uint64_t(p));
    oj[t] = int32_t(a[t].key % uint64_t(p));
    if (ov) ov[t] = a[t].v;
  }
}

}  // namespace cm
