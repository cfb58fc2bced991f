// atx_libm.h -- the host C library's sine, restated for the device.
//
// RFCavityPass evaluates one sine per particle and turn
// (atintegrators/attrackfunc.c:24-36). The reference calls the host libm,
// glibc 2.35 on the reference's platform (Ubuntu 22.04, x86-64), whose sin
// is the IBM Accurate Mathematical Library routine
// (sysdeps/ieee754/dbl-64/s_sin.c, as revised in glibc 2.28) built in the
// FMA variant that glibc's multiarch dispatch selects on hosts with FMA and
// AVX2. That routine is not correctly rounded (up to 0.548 ulp), so a
// correctly rounded device sine still differs from it in about one call in a
// thousand; over 1,000 turns such 1-ulp differences in delta decide the loss
// record of particles near the dynamic-aperture boundary. This file restates
// the routine's published algorithm, with the operations that the FMA build
// fuses written as fma(), so that the device returns the host's value bit
// for bit:
//
//   |x| < 2^-26                 sin x = x
//   |x| < 0.85546875            do_sin(x, 0): odd Taylor polynomial below
//                               0.126, else the table of sin/cos at k/128
//                               plus short polynomials for the remainder
//   |x| < 2.4262647628784180    cos(pi/2 - |x|) by the table (do_cos),
//                               pi/2 carried as two doubles
//   |x| < 105414350             Cody-Waite reduction by pi/2 in four
//                               pieces (n, a + da), then do_sin or do_cos
//   larger |x|, inf, nan        the device library's sin (glibc uses a
//                               Payne-Hanek reduction there; an RF phase of
//                               1e8 rad means |ct| > 1e7 m, never tracked)
//
// The table (atx_sincos_table.inc) is sin(k/128), cos(k/128) as
// head + tail doubles, computed from the series in exact decimal arithmetic
// by gen_sincos_table.py. The polynomial coefficients are the library's
// published ones. Every function is compiled without FP contraction
// whatever the translation unit's setting, so both kernel builds get the
// same bits; the host test (tests/test_libm_sin.py) compiles this header for
// the CPU and compares it with the host libm.
#ifndef ATX_LIBM_H
#define ATX_LIBM_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace atx {
namespace libm {

constexpr double kSinCos[128 * 4] = {
#include "atx_sincos_table.inc"
};

// sin x - x ~ x^3 (sn3 + x^2 sn5), 1 - cos x ~ x^2 (cs2 + x^2 (cs4 + x^2 cs6))
// for |x| <= 1/256
constexpr double kSn3 = -1.66666666666664880952546298448555E-01;
constexpr double kSn5 = 8.33333214285722277379541354343671E-03;
constexpr double kCs2 = 4.99999999999999999999950396842453E-01;
constexpr double kCs4 = -4.16666666666664434524222570944589E-02;
constexpr double kCs6 = 1.38888874007937613028114285595617E-03;
// sin x ~ x + x^3 (s1 + x^2 (s2 + x^2 (s3 + x^2 (s4 + x^2 s5)))) for |x| < 0.126
constexpr double kS1 = -0x1.5555555555555p-3;
constexpr double kS2 = 0x1.1111111110ECEp-7;
constexpr double kS3 = -0.00019841269834414642;
constexpr double kS4 = 2.755729806860771e-06;
constexpr double kS5 = -2.5022014848318398e-08;
// 1.5 * 2^45: |x| + kBig rounds |x| to a multiple of 1/128 (its index in the
// low word)
constexpr double kBig = 0x1.8p45;
// pi/2 = kHp0 + kHp1
constexpr double kHp0 = 0x1.921fb54442d18p0;
constexpr double kHp1 = 0x1.1a62633145c07p-54;
// Cody-Waite: pi/2 = kMp1 + kMp2 + kPp3 + kPp4 (leading pieces short enough
// that n * piece is exact), 2/pi and 1.5 * 2^52 for rounding to an integer
constexpr double kMp1 = 0x1.921fb58000000p0;
constexpr double kMp2 = -0x1.dde973c000000p-27;
constexpr double kPp3 = -0x1.cb3b398000000p-55;
constexpr double kPp4 = -0x1.d747f23e32ed7p-83;
constexpr double kHpInv = 0x1.45f306dc9c883p-1;
constexpr double kToInt = 0x1.8p52;

__host__ __device__ inline uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }

// |x| = (u - kBig) + rem with u - kBig = k/128; returns k
__host__ __device__ inline int table_split(double ax, double* rem) {
#pragma clang fp contract(off)
  const double u = kBig + ax;
  *rem = ax - (u - kBig);
  return (int)(bits(u) & 127u);
}

// sin(a + da) for |a| < 0.126
__host__ __device__ inline double taylor_sin(double xx, double a, double da) {
#pragma clang fp contract(off)
  const double p = __builtin_fma(
      __builtin_fma(__builtin_fma(__builtin_fma(kS5, xx, kS4), xx, kS3), xx, kS2), xx, kS1);
  const double t = __builtin_fma(__builtin_fma(p, a, -(0.5 * da)), xx, da);
  return a + t;
}

// sin(x + dx), |x| < 0.85546875
__host__ __device__ inline double do_sin(double x, double dx) {
#pragma clang fp contract(off)
  const double xold = x;
  if (__builtin_fabs(x) < 0.126) return taylor_sin(x * x, x, dx);
  if (x <= 0) dx = -dx;
  double r;
  const int k = table_split(__builtin_fabs(x), &r);
  const double xx = r * r;
  const double s = r + __builtin_fma(r * xx, __builtin_fma(xx, kSn5, kSn3), dx);
  const double c = __builtin_fma(r, dx, xx * __builtin_fma(xx, __builtin_fma(xx, kCs6, kCs4), kCs2));
  const double* t = kSinCos + 4 * k;
  const double cor = __builtin_fma(t[2], s, __builtin_fma(-t[0], c, __builtin_fma(s, t[3], t[1])));
  return __builtin_copysign(t[0] + cor, xold);
}

// cos(x + dx), |x| < 0.85546875
__host__ __device__ inline double do_cos(double x, double dx) {
#pragma clang fp contract(off)
  if (x < 0) dx = -dx;
  double r;
  const int k = table_split(__builtin_fabs(x), &r);
  r = r + dx;
  const double xx = r * r;
  const double s = __builtin_fma(r * xx, __builtin_fma(xx, kSn5, kSn3), r);
  const double c = xx * __builtin_fma(xx, __builtin_fma(xx, kCs6, kCs4), kCs2);
  const double* t = kSinCos + 4 * k;
  const double cor = __builtin_fma(-t[0], s, __builtin_fma(-t[2], c, __builtin_fma(-s, t[1], t[3])));
  return t[2] + cor;
}

// x = n pi/2 + (a + da), |a| <= pi/4 (up to rounding); returns n mod 4
__host__ __device__ inline int reduce_half_pi(double x, double* a, double* da) {
#pragma clang fp contract(off)
  const double t = __builtin_fma(x, kHpInv, kToInt);
  const double xn = t - kToInt;
  const double y = __builtin_fma(-xn, kMp2, __builtin_fma(-xn, kMp1, x));
  double t1 = xn * kPp3;
  const double t2 = y - t1;
  double db = (y - t2) - t1;
  t1 = xn * kPp4;
  const double b = t2 - t1;
  db += (t2 - b) - t1;
  *a = b;
  *da = db;
  return (int)(bits(t) & 3u);
}

__host__ __device__ inline double host_sin(double x) {
#pragma clang fp contract(off)
  const uint32_t k = (uint32_t)(bits(x) >> 32) & 0x7fffffffu;
  if (k < 0x3e500000u) return x;
  if (k < 0x3feb6000u) return do_sin(x, 0.0);
  if (k < 0x400368fdu) return __builtin_copysign(do_cos(kHp0 - __builtin_fabs(x), kHp1), x);
  if (k < 0x419921fbu) {
    double a, da;
    const int n = reduce_half_pi(x, &a, &da);
    const double v = (n & 1) ? do_cos(a, da) : do_sin(a, da);
    return (n & 2) ? -v : v;
  }
  return sin(x);
}

}  // namespace libm
}  // namespace atx

#endif  // ATX_LIBM_H
