// tests_runner: the C++ test executable of compmatmul_amd, in the role of the
// reference's Catch2 `tests` binary (tests.cpp).  Its known-answer checks
// (random streams, drawn hash parameters, hash values in closed form, small
// exact products, 16 x 16 products against GEMM) run against this library:
// the [cpu] cases against the host parameter drawing and the .npy I/O of the
// multiply tool, the [gpu] cases through the C-ABI on the device.
//
//   tests_runner            every case
//   tests_runner [cpu]      cases with the tag (also [gpu], [hash], ...)
//   tests_runner name       the case with that name
// Exit status: the number of failed cases (0 = all passed).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../compmatmul_amd/cli/npy.hpp"
#include "../../compmatmul_amd/csrc/host_hash.hpp"
#include "../../include/compmatmul_amd.h"

namespace {

struct Case {
  std::string name, tags;
  std::function<void()> fn;
};
std::vector<Case>& cases() {
  static std::vector<Case> v;
  return v;
}
struct Reg {
  Reg(const char* n, const char* t, std::function<void()> f) { cases().push_back({n, t, std::move(f)}); }
};
struct Failure : std::runtime_error {
  using std::runtime_error::runtime_error;
};
long g_checks = 0;

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define TEST_CASE(name, tags)                                    \
  static void CAT(case_, __LINE__)();                            \
  static Reg CAT(reg_, __LINE__)(name, tags, CAT(case_, __LINE__)); \
  static void CAT(case_, __LINE__)()
#define REQUIRE(expr)                                                                   \
  do {                                                                                  \
    ++g_checks;                                                                         \
    if (!(expr))                                                                        \
      throw Failure(std::string(__FILE__) + ":" + std::to_string(__LINE__) + ": " #expr); \
  } while (0)
#define CHECK_OK(call)                                                              \
  do {                                                                              \
    const int rc_ = (call);                                                         \
    if (rc_ != CM_OK)                                                               \
      throw Failure(std::string(#call) + " returned " + std::to_string(rc_) + ": " + \
                    cm_last_error());                                               \
  } while (0)

// xoshiro256+ doubles with splitmix64 seeding (Blackman and Vigna), the
// generator of the reference's drng (rng.hpp:86-107)
struct Drng {
  uint64_t s[4];
  explicit Drng(uint64_t seed) {
    for (auto& w : s) {
      uint64_t z = (seed += 0x9e3779b97f4a7c15ull);
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      w = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  double operator()() {
    const uint64_t r = s[0] + s[3];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return double(r >> 11) * 0x1.0p-53;
  }
};

bool almost_equal(size_t n, const double* a, const double* b, double tol = 1e-6) {
  for (size_t i = 0; i < n; ++i)
    if (std::fabs(a[i] - b[i]) > tol) return false;
  return true;
}

// small exact products of the reference tests (tests.cpp:37-88)
std::vector<double> iota_d(double from, int n) {
  std::vector<double> v(n);
  for (int i = 0; i < n; ++i) v[i] = from + i;
  return v;
}
std::vector<double> transpose(const std::vector<double>& a, int rows, int cols) {
  std::vector<double> t(a.size());
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) t[size_t(j) * rows + i] = a[size_t(i) * cols + j];
  return t;
}
struct Small {
  int m, n, p;
  std::vector<double> AT, B, C;
};
std::vector<Small> small_products() {
  // A is m x n row-major with entries 1.., AT its transpose; B n x p
  auto make = [](int m, int n, int p, double b0, std::vector<double> C) {
    Small s{m, n, p, transpose(iota_d(1, m * n), m, n), iota_d(b0, n * p), std::move(C)};
    return s;
  };
  return {
      make(4, 4, 4, 17, {250, 260, 270, 280, 618, 644, 670, 696, 986, 1028, 1070, 1112, 1354, 1412, 1470, 1528}),
      make(4, 2, 4, 9, {35, 38, 41, 44, 79, 86, 93, 100, 123, 134, 145, 156, 167, 182, 197, 212}),
      make(2, 4, 2, 9, {130, 140, 322, 348}),
      make(2, 4, 8, 1, {170, 180, 190, 200, 210, 220, 230, 240, 378, 404, 430, 456, 482, 508, 534, 560}),
  };
}

std::vector<uint64_t> draw(int family, int n, uint64_t seed) {
  std::vector<uint64_t> out;
  cm::draw_multiply(family, n, seed, out);
  return out;
}

// ------------------------------------------------------------------ [cpu]
TEST_CASE("test_rng", "[rng][cpu]") {  // tests.cpp:90-96
  cm::Xoshiro256ss rng(1234);
  REQUIRE(rng() == 0x0bab45d9a0e3ae53ull);
  REQUIRE(rng() == 0xd7c640660c19433eull);
  REQUIRE(rng() == 0xb0dedaa0d09a6691ull);
  REQUIRE(rng() == 0xdec9f41b58ec86ebull);
}

TEST_CASE("test_drawn_parameters", "[hash][cpu]") {
  // multiply-shift draw_one(64, 123456): odd rejection (tests.cpp:146-160)
  REQUIRE(draw(cm::kMultiplyShift, 1, 123456)[0] == 2749059519329956733ull);
  // multiply-shift draw(2, 32, 19850) (tests.cpp:185-186)
  auto ms = draw(cm::kMultiplyShift, 2, 19850);
  REQUIRE(ms[0] == 3688180003657674949ull && ms[1] == 11729574476469721337ull);
  // multiply-add-shift draw(2, 128, 22201) (tests.cpp:272-275)
  auto mas = draw(cm::kMultiplyAddShift, 2, 22201);
  REQUIRE(mas[0] == 2016826348251915964ull && mas[1] == 17696970887482165528ull);
  REQUIRE(mas[2] == 6666994550483350943ull && mas[3] == 1097363203091539945ull);
  // multiply-add-shift draw_one(32, 8876) (tests.cpp:797-801)
  auto one = draw(cm::kMultiplyAddShift, 1, 8876);
  REQUIRE(one[0] == 15506110424104283675ull && one[1] == 3138147583905456810ull);
  // sketch(d = 2, b = 2, seed 3483): h1, h2, s1, s2 from seed + 0..3 (tests.cpp:819-840)
  const uint64_t want[4][4] = {
      {1312110179027324775ull, 15465668779316680995ull, 4555306672462241870ull, 12288159137770073612ull},
      {4329212686603688483ull, 10307990599335447505ull, 12597734872139453857ull, 15119207240207437602ull},
      {11899074044546628492ull, 298345645907162146ull, 8628211646415106536ull, 12647867801867038929ull},
      {7008367753992221008ull, 16948937881964684143ull, 10319872211339334476ull, 13096518497071785733ull}};
  for (int f = 0; f < 4; ++f) {
    auto p = draw(cm::kMultiplyAddShift, 2, 3483 + f);
    for (int k = 0; k < 4; ++k) REQUIRE(p[k] == want[f][k]);
  }
}

TEST_CASE("test_tabulation_tables", "[hash][cpu]") {  // tests.cpp:312-320, 349-353
  for (int family : {int(cm::kTab8), int(cm::kTab16)}) {
    const uint64_t seed = family == cm::kTab8 ? 23711 : 23712;
    const size_t entries = family == cm::kTab8 ? 4 * 256 : 2 * 65536;
    std::vector<uint32_t> tabs;
    cm::draw_tables(family, 1, seed, tabs);
    REQUIRE(tabs.size() == entries);
    cm::Xoshiro256ss prng(seed);  // row-major fill, low 32 bits of each draw
    for (size_t k = 0; k < entries; ++k) REQUIRE(tabs[k] == uint32_t(prng()));
  }
}

TEST_CASE("test_npy_round_trip", "[npy][cpu]") {
  const std::string path = "/tmp/cm_runner_rt.npy";
  std::vector<double> v(6 * 7);
  for (size_t i = 0; i < v.size(); ++i) v[i] = 0.5 * double(i) - 3.0;
  cmnpy::save_f64(path, v.data(), {6, 7});
  cmnpy::Array a = cmnpy::load(path);
  REQUIRE(a.descr == "<f8" && a.word_size == 8 && !a.fortran_order);
  REQUIRE(a.shape.size() == 2 && a.shape[0] == 6 && a.shape[1] == 7);
  REQUIRE(std::memcmp(a.data.data(), v.data(), v.size() * 8) == 0);
  // a version-2 header in Fortran order with double quotes, written by hand
  std::string dict = "{\"descr\": \"<f4\", \"fortran_order\": True, \"shape\": (3,), }";
  dict.append((64 - (12 + dict.size() + 1) % 64) % 64, ' ');
  dict += '\n';
  FILE* f = std::fopen(path.c_str(), "wb");
  REQUIRE(f != nullptr);
  std::fwrite("\x93NUMPY\x02\x00", 1, 8, f);
  const uint32_t hl = uint32_t(dict.size());
  const unsigned char lb[4] = {uint8_t(hl), uint8_t(hl >> 8), uint8_t(hl >> 16), uint8_t(hl >> 24)};
  std::fwrite(lb, 1, 4, f);
  std::fwrite(dict.data(), 1, dict.size(), f);
  const float fv[3] = {1.f, 2.f, 3.f};
  std::fwrite(fv, 4, 3, f);
  std::fclose(f);
  cmnpy::Array b = cmnpy::load(path);
  REQUIRE(b.descr == "<f4" && b.word_size == 4 && b.fortran_order);
  REQUIRE(b.shape.size() == 1 && b.shape[0] == 3 && b.as<float>()[2] == 3.f);
  std::remove(path.c_str());
}

// ------------------------------------------------------------------ [gpu]
// packed entries: bucket in bits 0..30, bit 31 set for sign -1
uint32_t bucket(uint32_t e) { return e & 0x7fffffffu; }
int sign(uint32_t e) { return (e >> 31) ? -1 : 1; }

TEST_CASE("test_multiply_shift", "[hash][gpu]") {  // tests.cpp:146-225
  const uint64_t a = 2749059519329956733ull;
  const int n = 129;
  std::vector<uint32_t> hA(n), hB(1);
  // h1 is drawn from the seed itself, s1 from seed + 2
  CHECK_OK(cm_hash_arrays(CM_MULTIPLY_SHIFT, 1, 64, 123456, n, 1, hA.data(), hB.data(), CM_HOST));
  for (uint64_t i = 0; i < uint64_t(n); ++i) REQUIRE(bucket(hA[i]) == uint32_t((a * i) >> 58));
  CHECK_OK(cm_hash_arrays(CM_MULTIPLY_SHIFT, 1, 64, 123456 - 2, n, 1, hA.data(), hB.data(), CM_HOST));
  for (uint64_t i = 0; i < uint64_t(n); ++i) REQUIRE(sign(hA[i]) == (((a * i) >> 63) == 1 ? 1 : -1));
  // two functions drawn from 19850, b = 32
  const uint64_t a1 = 3688180003657674949ull, a2 = 11729574476469721337ull;
  std::vector<uint32_t> h2(2 * 1025), hb(2);
  CHECK_OK(cm_hash_arrays(CM_MULTIPLY_SHIFT, 2, 32, 19850, 1025, 1, h2.data(), hb.data(), CM_HOST));
  for (uint64_t i = 0; i <= 1024; ++i) {
    REQUIRE(bucket(h2[i]) == uint32_t((a1 * i) >> 59));
    REQUIRE(bucket(h2[1025 + i]) == uint32_t((a2 * i) >> 59));
  }
}

TEST_CASE("test_multiply_add_shift", "[hash][gpu]") {  // tests.cpp:228-309
  const uint64_t a = 2366053268901514180ull, b = 2749059519329956733ull;
  const int n = 513;
  std::vector<uint32_t> hA(n), hB(1);
  CHECK_OK(cm_hash_arrays(CM_MULTIPLY_ADD_SHIFT, 1, 128, 123456, n, 1, hA.data(), hB.data(), CM_HOST));
  for (uint64_t i = 0; i < uint64_t(n); ++i) REQUIRE(bucket(hA[i]) == uint32_t((a * i + b) >> 57));
  CHECK_OK(cm_hash_arrays(CM_MULTIPLY_ADD_SHIFT, 1, 128, 123456 - 2, n, 1, hA.data(), hB.data(), CM_HOST));
  for (uint64_t i = 0; i < uint64_t(n); ++i) REQUIRE(sign(hA[i]) == (((a * i + b) >> 63) == 1 ? 1 : -1));
  const uint64_t a1 = 2016826348251915964ull, b1 = 17696970887482165528ull;
  const uint64_t a2 = 6666994550483350943ull, b2 = 1097363203091539945ull;
  std::vector<uint32_t> h2(2 * 2049), hb(2);
  CHECK_OK(cm_hash_arrays(CM_MULTIPLY_ADD_SHIFT, 2, 128, 22201, 2049, 1, h2.data(), hb.data(), CM_HOST));
  for (uint64_t i = 0; i <= 2048; ++i) {
    REQUIRE(bucket(h2[i]) == uint32_t((a1 * i + b1) >> 57));
    REQUIRE(bucket(h2[2049 + i]) == uint32_t((a2 * i + b2) >> 57));
  }
}

TEST_CASE("test_tabulation_hash", "[hash][gpu]") {  // tests.cpp:312-460
  for (int family : {CM_TABULATION8, CM_TABULATION16}) {
    const uint64_t seed = 23711;
    std::vector<uint32_t> T;
    cm::draw_tables(family, 1, seed, T);
    const int n = 70000;  // keys past 2^16 exercise every character table
    std::vector<uint32_t> hA(n), hB(1);
    CHECK_OK(cm_hash_arrays(family, 1, 256, seed, n, 1, hA.data(), hB.data(), CM_HOST));
    std::vector<uint32_t> S;
    cm::draw_tables(family, 1, seed + 2, S);
    for (uint32_t x = 0; x < uint32_t(n); ++x) {
      uint32_t h, s;
      if (family == CM_TABULATION8) {
        h = T[x & 0xff] ^ T[256 + ((x >> 8) & 0xff)] ^ T[512 + ((x >> 16) & 0xff)] ^ T[768 + (x >> 24)];
        s = S[x & 0xff] ^ S[256 + ((x >> 8) & 0xff)] ^ S[512 + ((x >> 16) & 0xff)] ^ S[768 + (x >> 24)];
      } else {
        h = T[x & 0xffff] ^ T[65536 + (x >> 16)];
        s = S[x & 0xffff] ^ S[65536 + (x >> 16)];
      }
      REQUIRE(bucket(hA[x]) == (h >> 24));
      REQUIRE(sign(hA[x]) == ((s >> 31) ? 1 : -1));
    }
  }
}

TEST_CASE("test_gemm", "[matmul][gpu]") {  // tests.cpp:722-750
  for (const Small& s : small_products()) {
    std::vector<double> C(size_t(s.m) * s.p, -1.0);
    CHECK_OK(cm_gemm(s.m, s.n, s.p, C.data(), s.AT.data(), s.B.data()));
    REQUIRE(C == s.C);
  }
}

void small_compressed(bool fwht, const std::vector<std::pair<int, int>>& dbs) {
  for (auto [d, b] : dbs) {  // tests.cpp:844-936 / :1052-1145
    if (!(d > 1 && b >= 32)) continue;  // the reference checks values only then
    for (const Small& s : small_products()) {
      std::vector<double> C(size_t(s.m) * s.p, 0.0);
      const int rc = fwht ? cm_matmul_fwht(s.m, s.n, s.p, C.data(), s.AT.data(), s.B.data(), d, b, 25926, 1, CM_MULTIPLY_ADD_SHIFT)
                          : cm_matmul_fft(s.m, s.n, s.p, C.data(), s.AT.data(), s.B.data(), d, b, 25926, 1, CM_MULTIPLY_ADD_SHIFT);
      CHECK_OK(rc);
      REQUIRE(almost_equal(C.size(), C.data(), s.C.data()));
    }
  }
}

void sixteen(bool fwht) {  // tests.cpp:940-1006 / :1148-1214
  const int n = 16;
  const uint64_t seed = 15954;
  Drng drng(seed);
  std::vector<double> AT(n * n), B(n * n), C1(n * n);
  for (int i = 0; i < n * n; ++i) {
    AT[i] = drng();
    B[i] = drng();
  }
  CHECK_OK(cm_gemm(n, n, n, C1.data(), AT.data(), B.data()));
  struct P {
    int family, d, b;
  };
  const P fft[4] = {{CM_MULTIPLY_ADD_SHIFT, 5, 1024}, {CM_MULTIPLY_SHIFT, 5, 1024},
                    {CM_TABULATION8, 11, 2048}, {CM_TABULATION16, 11, 1024}};
  const P wht[4] = {{CM_MULTIPLY_ADD_SHIFT, 5, 2048}, {CM_MULTIPLY_SHIFT, 5, 2048},
                    {CM_TABULATION8, 11, 2048}, {CM_TABULATION16, 11, 1024}};
  for (int k = 0; k < 4; ++k) {
    const P& q = fwht ? wht[k] : fft[k];
    std::vector<double> C(n * n);
    const int rc = fwht ? cm_matmul_fwht(n, n, n, C.data(), AT.data(), B.data(), q.d, q.b, seed + 1 + k, 1, q.family)
                        : cm_matmul_fft(n, n, n, C.data(), AT.data(), B.data(), q.d, q.b, seed + 1 + k, 1, q.family);
    CHECK_OK(rc);
    REQUIRE(almost_equal(C.size(), C.data(), C1.data(), 1e-12));
  }
}

TEST_CASE("test_matmul_fft", "[matmul][gpu]") {
  small_compressed(false, {{1, 2}, {1, 8}, {1, 32}, {3, 2}, {3, 8}, {5, 32}});
  sixteen(false);
}

TEST_CASE("test_matmul_fwht", "[matmul][gpu]") {
  small_compressed(true, {{1, 2}, {1, 8}, {1, 32}, {3, 2}, {3, 8}, {7, 32}});
  sixteen(true);
}

bool selected(const Case& c, const std::vector<std::string>& filters) {
  if (filters.empty()) return true;
  for (const std::string& f : filters) {
    if (f.size() > 2 && f.front() == '[' && c.tags.find(f) != std::string::npos) return true;
    if (f == c.name) return true;
  }
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> filters(argv + 1, argv + argc);
  int ran = 0, failed = 0;
  for (const Case& c : cases()) {
    if (!selected(c, filters)) continue;
    ++ran;
    try {
      c.fn();
      std::printf("PASSED %s %s\n", c.name.c_str(), c.tags.c_str());
    } catch (const std::exception& e) {
      ++failed;
      std::printf("FAILED %s %s\n  %s\n", c.name.c_str(), c.tags.c_str(), e.what());
    }
    std::fflush(stdout);
  }
  std::printf("%d test cases, %d failed, %ld checks\n", ran, failed, g_checks);
  return failed;
}
