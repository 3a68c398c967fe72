"""The parallel host copy of the transfer path (csrc/host_copy.cpp), built on
its own with g++: copies of sizes around the parallel threshold and not
multiples of the part size arrive intact; eight threads copying at once (one
runs on the pool, the others copy alone) all arrive intact; a forked child
copies with a pool of its own."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "compmatmul_amd", "csrc")

PROGRAM = r'''
#include <sys/wait.h>
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#include "host_copy.hpp"

static bool check(size_t len, unsigned seed) {
  std::vector<unsigned char> src(len + 3), dst(len + 3, 0xEE);
  for (size_t i = 0; i < src.size(); ++i) src[i] = (unsigned char)((i * 131 + seed * 7) >> 3);
  // unaligned start, so parts do not begin on 64-byte boundaries of the buffers
  cm::parallel_copy(dst.data() + 1, src.data() + 1, len);
  if (dst[0] != 0xEE || dst[len + 1] != 0xEE || dst[len + 2] != 0xEE) return false;
  return std::memcmp(dst.data() + 1, src.data() + 1, len) == 0;
}

int main() {
  const size_t sizes[] = {0, 1, 63, 4096, (512u << 10) - 1, 512u << 10, (512u << 10) + 1,
                          (1u << 20) + 3, (7u << 20) + 13, 33554432};
  for (size_t s : sizes)
    if (!check(s, unsigned(s % 97))) { std::printf("size %zu failed\n", s); return 1; }
  // concurrent callers
  std::vector<std::thread> ts;
  std::vector<int> ok(8, 0);
  for (int t = 0; t < 8; ++t)
    ts.emplace_back([t, &ok] {
      bool good = true;
      for (int r = 0; r < 20; ++r) good = good && check((3u << 20) + size_t(t) * 4099 + r, t * 31 + r);
      ok[t] = good;
    });
  for (auto& th : ts) th.join();
  for (int t = 0; t < 8; ++t)
    if (!ok[t]) { std::printf("thread %d failed\n", t); return 1; }
  // a forked child (the pool's threads do not exist there)
  const pid_t pid = fork();
  if (pid == 0) _exit(check(5u << 20, 3) && check((2u << 20) + 7, 4) ? 0 : 1);
  int status = 0;
  waitpid(pid, &status, 0);
  if (!WIFEXITED(status) || WEXITSTATUS(status) != 0) { std::printf("child failed\n"); return 1; }
  if (!check(9u << 20, 5)) { std::printf("after fork failed\n"); return 1; }
  std::printf("ok\n");
  return 0;
}
'''


def test_parallel_copy(tmp_path):
    src, exe = tmp_path / "c.cpp", tmp_path / "c"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", CSRC, str(src),
                    os.path.join(CSRC, "host_copy.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
