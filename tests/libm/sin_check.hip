// Host check of atx_libm.h: compares atx::libm::host_sin with the host C
// library's sin, bit for bit, on seeded random arguments in [lo, hi) (both
// signs) and on the edge arguments of every branch.
//   sin_check LO HI N SEED  -> prints "checked M mismatches K" and the first few
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../at_amd/csrc/atx_libm.h"

static double from_bits(uint64_t b) {
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}

static long mismatches = 0, checked = 0;

static void check(double x) {
  const double a = std::sin(x), b = atx::libm::host_sin(x);
  checked++;
  if (std::memcmp(&a, &b, 8) != 0 && !(std::isnan(a) && std::isnan(b))) {
    if (mismatches < 5) std::printf("x=%a libm=%a restated=%a\n", x, a, b);
    mismatches++;
  }
}

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  const double lo = std::atof(argv[1]), hi = std::atof(argv[2]);
  const long n = std::atol(argv[3]);
  std::mt19937_64 gen(std::strtoull(argv[4], nullptr, 10));
  std::uniform_real_distribution<double> u(lo, hi);
  for (long i = 0; i < n; i++) check((i & 1) ? -u(gen) : u(gen));
  // branch edges (high words of the thresholds), +-0, table ties k/256
  const uint64_t edges[] = {0x3e500000ull << 32, 0x3feb6000ull << 32, 0x400368fdull << 32,
                            0x419921fbull << 32, 0x3fc020c49ba5e354ull /* 0.126 */};
  for (uint64_t e : edges)
    for (int d = -3; d <= 3; d++) {
      check(from_bits(e + d));
      check(-from_bits(e + d));
    }
  check(0.0);
  check(-0.0);
  for (int k = 0; k < 256; k++) {
    check(k / 256.0);
    check(-k / 256.0);
    check(std::nextafter(k / 256.0, 1.0));
  }
  std::printf("checked %ld mismatches %ld\n", checked, mismatches);
  return 0;
}
