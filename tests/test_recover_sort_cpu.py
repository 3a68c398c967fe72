"""The host ordering of recovered entries (csrc/recover_sort.hpp) against
numpy's lexicographic sort by (i, j), on random lists: small and large
column counts (keys past 32 bits), lists of every length up to a few
thousand, values carried with their entries."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r'''
#include <cstdio>
#include <vector>
#include "recover_sort.hpp"
int main() {
  long long n, p;
  while (std::scanf("%lld %lld", &n, &p) == 2) {
    std::vector<int32_t> oi(n), oj(n);
    std::vector<float> ov(n);
    for (long long t = 0; t < n; ++t) std::scanf("%d %d %f", &oi[t], &oj[t], &ov[t]);
    cm::sort_recovered<float>(n, p, oi.data(), oj.data(), ov.data());
    for (long long t = 0; t < n; ++t) std::printf("%d %d %.9g\n", oi[t], oj[t], ov[t]);
  }
}
'''


@pytest.fixture(scope="module")
def sorter(tmp_path_factory):
    d = tmp_path_factory.mktemp("sort")
    src, exe = d / "s.cpp", d / "s"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "compmatmul_amd", "csrc"),
                    str(src), "-o", str(exe)], check=True)
    return str(exe)


@pytest.mark.parametrize("m,p", [(1, 1), (64, 64), (4096, 4096), (3, 1 << 30), (1 << 20, 1 << 20)])
def test_sort_matches_lexicographic_order(sorter, m, p):
    rng = np.random.default_rng(m + p)
    cases = []
    for n in (0, 1, 2, 3, 255, 256, 257, 4000):
        n = min(n, m * p)
        flat = rng.choice(m * p, size=n, replace=False) if m * p < 1 << 24 else \
            np.unique(rng.integers(0, m * p, size=n, dtype=np.int64))
        rng.shuffle(flat)
        oi, oj = (flat // p).astype(np.int64), (flat % p).astype(np.int64)
        ov = rng.standard_normal(len(flat)).astype(np.float32)
        cases.append((oi, oj, ov))
    text = "".join(f"{len(oi)} {p}\n" + "".join(f"{a} {b} {c:.9g}\n" for a, b, c in zip(oi, oj, ov))
                   for oi, oj, ov in cases)
    out = subprocess.run([sorter], input=text, capture_output=True, text=True, check=True).stdout.split()
    pos = 0
    for oi, oj, ov in cases:
        k = len(oi)
        got = np.array(out[pos:pos + 3 * k], dtype=np.float64).reshape(k, 3)
        pos += 3 * k
        order = np.lexsort((oj, oi))
        assert np.array_equal(got[:, 0], oi[order])
        assert np.array_equal(got[:, 1], oj[order])
        assert np.array_equal(got[:, 2].astype(np.float32), ov[order])
