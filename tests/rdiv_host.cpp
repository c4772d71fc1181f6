// Host run of the corrected-reciprocal division of bio-cfd_amd/csrc/rdiv.hpp
// (the momentum kernel's divisions by the upwind denominators and re)
// against the division x / b, bit for bit: random significands and
// exponents over the range the kernel uses it in (|x| from 2^-956 to 2^699,
// b from 2^-40 to 2^40), quotients next to rounding midpoints (x = b q with
// q one ulp off a representable value), divisors whose significand is all
// ones or nearly 1, zero numerators of both signs, and the kernel's own
// denominators (ca6 h for grids of 1/8 to 1/128 and upwind tables of
// stretched grids).
//
//   rdiv_host N        exit 0 when all N cases (plus the fixed ones) agree
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "rdiv.hpp"

using bcfd::rdiv_core;

static uint64_t g_s = 0x9e3779b97f4a7c15ull;
static uint64_t rnd() {
  g_s ^= g_s << 13;
  g_s ^= g_s >> 7;
  g_s ^= g_s << 17;
  return g_s;
}
static double from_bits(uint64_t u) {
  double d;
  memcpy(&d, &u, 8);
  return d;
}
static uint64_t bits(double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  return u;
}
static double rnd_double(int emin, int emax) {
  const int e = emin + (int)(rnd() % (uint64_t)(emax - emin + 1));
  return from_bits((rnd() & 0x000FFFFFFFFFFFFFull) | ((uint64_t)(1023 + e) << 52)) * ((rnd() & 1) ? -1.0 : 1.0);
}

static long g_bad = 0, g_n = 0;
static void check(double x, double b) {
  const double y = 1.0 / b;
  const double ref = x / b, got = rdiv_core(x, b, y);
  ++g_n;
  if (bits(ref) != bits(got)) {
    if (g_bad < 10) printf("mismatch: x=%a b=%a x/b=%a got=%a\n", x, b, ref, got);
    ++g_bad;
  }
}

int main(int argc, char **argv) {
  const long n = argc > 1 ? atol(argv[1]) : 20000000;
  for (long t = 0; t < n; ++t) {
    const double b = std::fabs(rnd_double(-40, 40));
    switch (t % 4) {
      case 0: check(rnd_double(-956, 699), b); break;
      case 1: {  // next to a midpoint: x = b q with q one ulp off a double
        const double q = rnd_double(-300, 300);
        check(b * std::nextafter(q, 2 * q), b);
        break;
      }
      case 2: {  // divisor significand all ones or nearly 1
        const uint64_t e = (uint64_t)(1023 + (int)(rnd() % 41) - 20) << 52;
        const uint64_t m = (rnd() & 1) ? (0x000FFFFFFFFFFFFFull ^ (rnd() & 0xFF)) : (rnd() & 0xFF);
        check(rnd_double(-600, 600), from_bits(e | m));
        break;
      }
      default: check(rnd_double(-100, 100), b); break;
    }
  }
  // zero numerators keep their sign
  for (double b : {0.375, 3.0, 1e-10, 7.5e8}) {
    check(0.0, b);
    check(-0.0, b);
  }
  // the corners of the range: numerators at the smallest and largest
  // exponents the kernel admits over divisors at 2^-40 and 2^40
  for (long t = 0; t < 400000; ++t) {
    const double x = (t & 1) ? rnd_double(-956, -940) : rnd_double(684, 699);
    const double b = std::ldexp(1.0 + (double)(rnd() % 4096) / 4096.0, (t & 2) ? 39 : -40);
    check(x, b);
    check(x, std::ldexp(1.0, (t & 2) ? 40 : -40));
  }
  // the kernel's denominators: ca6 h with third-order upwind coefficients of
  // stretched grids (ratios 0.8 .. 1.25) and spacings 1/8 .. 1/128
  for (long t = 0; t < 200000; ++t) {
    const double h = std::ldexp(1.0, -3 - (int)(rnd() % 5)) * (0.8 + 0.45 * (double)(rnd() % 1000) / 1000.0);
    const double c = 1.0 + 2.0 * (double)(rnd() % 100000) / 100000.0;
    check(rnd_double(-60, 20), c * h);
    check(rnd_double(-60, 20), 2.0 * c * h);
  }
  printf("%ld cases, %ld mismatches\n", g_n, g_bad);
  if (g_bad == 0) printf("every quotient equals the division\n");
  return g_bad == 0 ? 0 : 1;
}
