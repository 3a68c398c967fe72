// Driver that compiles the reference's own rng.hpp (read in place from the
// reference checkout, never copied) so the oracle's generator can be checked
// against it.  Usage: ref_rng <seed> <count>
// Prints <count> xoshiro256** words (hex) then <count> xoshiro256+ doubles.
#include "rng.hpp"
#include <cstdio>
#include <cstdlib>

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  const uint64_t seed = std::strtoull(argv[1], nullptr, 10);
  const long n = std::strtol(argv[2], nullptr, 10);
  compmatmul::rng r(seed);
  for (long i = 0; i < n; ++i) std::printf("%016llx\n", (unsigned long long)r());
  compmatmul::drng dr(seed);
  for (long i = 0; i < n; ++i) std::printf("%.17g\n", dr());
  return 0;
}
