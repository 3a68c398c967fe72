// Host-side random parameter drawing for the four hash families.
//
// Behaviour follows the reference bit for bit:
//   generator   rng.hpp:26-33 (splitmix64 seeding), :59-72 (xoshiro256**)
//   default     matmul_fft.hpp:30-32 / rng.hpp:40-44 (first draw of a
//               generator seeded with wall-clock milliseconds)
//   families    hashing.hpp:69-87 (multiply-shift, odd rejection),
//               :145-160 (multiply-add-shift, a then b), :277-295
//               (tabulation tables filled row-major, low 32 bits per draw)
//   sketch      matmul_sketch.hpp:15-18 (h1, h2, s1, s2 from seed+0..3)
// The device evaluates the drawn functions (hash.hip).
#pragma once
#include <cstdint>
#include <vector>

namespace cm {

enum Family : int { kMultiplyAddShift = 0, kMultiplyShift = 1, kTab8 = 2, kTab16 = 3 };

class Xoshiro256ss {
 public:
  explicit Xoshiro256ss(uint64_t seed);
  uint64_t operator()();

 private:
  uint64_t s_[4];
};

uint64_t time_seed();

// words of parameter data per function (u64 for the multiply families,
// u32 table entries stored one per u64 word for tabulation)
int64_t param_words(int family);

// Parameters of n functions drawn from one stream seeded with `seed`.
// Tabulation tables are returned as u32 (n * words entries).
void draw_multiply(int family, int n, uint64_t seed, std::vector<uint64_t>& out);
// The same parameters into out[0 .. n * param_words(family)); no allocation.
void draw_multiply_into(int family, int n, uint64_t seed, uint64_t* out);
void draw_tables(int family, int n, uint64_t seed, std::vector<uint32_t>& out);

}  // namespace cm
