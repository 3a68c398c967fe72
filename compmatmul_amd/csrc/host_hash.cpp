#include "host_hash.hpp"

#include <chrono>

namespace cm {

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

Xoshiro256ss::Xoshiro256ss(uint64_t x) {
  for (int i = 0; i < 4; ++i) {  // splitmix64
    x += 0x9e3779b97f4a7c15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    s_[i] = z ^ (z >> 31);
  }
}

uint64_t Xoshiro256ss::operator()() {
  const uint64_t out = rotl(s_[1] * 5, 7) * 9;
  const uint64_t t = s_[1] << 17;
  s_[2] ^= s_[0];
  s_[3] ^= s_[1];
  s_[1] ^= s_[2];
  s_[0] ^= s_[3];
  s_[2] ^= t;
  s_[3] = rotl(s_[3], 45);
  return out;
}

uint64_t time_seed() {
  using namespace std::chrono;
  const uint64_t ms =
      duration_cast<milliseconds>(system_clock::now().time_since_epoch()).count();
  Xoshiro256ss r(ms);
  return r();
}

int64_t param_words(int family) {
  switch (family) {
    case kMultiplyShift: return 1;
    case kMultiplyAddShift: return 2;
    case kTab8: return 4 * 256;
    case kTab16: return 2 * 65536;
    default: return -1;
  }
}

void draw_multiply_into(int family, int n, uint64_t seed, uint64_t* out) {
  Xoshiro256ss r(seed);
  for (int f = 0; f < n; ++f) {
    if (family == kMultiplyShift) {
      uint64_t a;
      do { a = r(); } while ((a & 1) == 0);
      *out++ = a;
    } else {
      const uint64_t a = r();
      const uint64_t b = r();
      *out++ = a;
      *out++ = b;
    }
  }
}

void draw_multiply(int family, int n, uint64_t seed, std::vector<uint64_t>& out) {
  out.resize(size_t(n) * size_t(param_words(family)));
  draw_multiply_into(family, n, seed, out.data());
}

void draw_tables(int family, int n, uint64_t seed, std::vector<uint32_t>& out) {
  Xoshiro256ss r(seed);
  const int64_t w = param_words(family);
  out.resize(size_t(n) * size_t(w));
  for (auto& v : out) v = uint32_t(r());
}

}  // namespace cm
