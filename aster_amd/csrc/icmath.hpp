// Random streams and elementary functions for the seeded initial-condition
// generators, shared by the host generators (ic.cpp) and the device
// generators (ic.hip) so that both produce the same bits for the same seed.
//
// Only +, -, *, / and sqrt (IEEE-correct on the host and on gfx950 in FP64),
// exact scalings by powers of two (ldexp, frexp), rint and integer
// conversions are used, and every translation unit that includes this header
// is compiled without floating-point contraction.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define ICM_HD __host__ __device__ inline
#else
#define ICM_HD inline
#endif

namespace aster {
namespace icm {

constexpr double kPi = 3.14159265358979323846;

// ln 2 split into a 32-bit head and a tail: k * kLn2Hi is exact for |k| < 2^21.
constexpr double kLn2Hi = 0x1.62e42fee00000p-1;
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kLn10 = 0x1.26bb1bbb55516p+1;
constexpr double kSqrtHalf = 0x1.6a09e667f3bcdp-1;
// pi/2 in three parts, the first two with 33 significant bits: q * part is
// exact for |q| < 2^20.
constexpr double kPio2_1 = 0x1.921fb54400000p+0;
constexpr double kPio2_2 = 0x1.0b4611a600000p-34;
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;
constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;

// SplitMix64 finaliser (Steele, Lea & Flood 2014; Vigna's splitmix64.c).
ICM_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Counter-based stream for item i of a seeded generator: the key is
// mix64(seed ^ mix64(i + C)); draw k (k = 1, 2, ...) is the k-th output of a
// SplitMix64 generator whose state starts at the key, i.e.
// mix64(key + k * 0x9E3779B97F4A7C15), and u() maps its top 53 bits to
// [0, 1). Items are independent, so any subset is generated in any order.
struct Stream {
  uint64_t key, k;
  ICM_HD Stream(uint64_t seed, uint64_t i) : key(mix64(seed ^ mix64(i + 0x632BE59BD9B4E019ull))), k(0) {}
  ICM_HD uint64_t next() {
    ++k;
    return mix64(key + k * 0x9E3779B97F4A7C15ull);
  }
  ICM_HD double u() { return (double)(next() >> 11) * 0x1.0p-53; }
  // (0, 1]: never zero.
  ICM_HD double u_open0() { return (double)((next() >> 11) + 1) * 0x1.0p-53; }
};

// e^x by Cody & Waite's reduction x = k ln2 + r (|r| <= ln2 / 2, ln 2 in
// head and tail), the Taylor series of e^r to degree 13 (truncation below
// 4e-18 relative) in Horner form, and an exact scaling by 2^k.
ICM_HD double exp(double x) {
  if (x != x) return x;
  if (x > 709.78) return HUGE_VAL;
  if (x < -745.2) return 0.0;
  const double k = rint(x * kInvLn2);
  const double r = (x - k * kLn2Hi) - k * kLn2Lo;
  double p = 1.0 / 6227020800.0;  // 1/13!
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

// ln x for x > 0: x = (1 + f) 2^e with 1 + f in [sqrt(1/2), sqrt(2)), and
// ln(1 + f) = 2 atanh(s), s = f / (2 + f), |s| <= 0.1716, arranged as in
// Tang's table-free method (the fdlibm log): 2 atanh(s) = 2s + s R with
// R = sum 2 s^(2k) / (2k + 1) (to s^22, truncation below 1e-19), and
// 2s = f - s f = f - (f^2/2 - s f^2/2), so
// ln(1 + f) = f - (f^2/2 - s (f^2/2 + R)): the exact f carries the result and
// only the small correction is rounded. Plus e ln2 in head and tail.
ICM_HD double log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -HUGE_VAL : (x - x) / (x - x);
  if (x == HUGE_VAL) return x;
  int e = 0;
  double m = frexp(x, &e);  // m in [0.5, 1)
  if (m < kSqrtHalf) {
    m = m * 2.0;
    e -= 1;
  }
  const double f = m - 1.0;  // exact
  const double s = f / (2.0 + f);
  const double s2 = s * s;
  double p = 2.0 / 23.0;
  p = p * s2 + 2.0 / 21.0;
  p = p * s2 + 2.0 / 19.0;
  p = p * s2 + 2.0 / 17.0;
  p = p * s2 + 2.0 / 15.0;
  p = p * s2 + 2.0 / 13.0;
  p = p * s2 + 2.0 / 11.0;
  p = p * s2 + 2.0 / 9.0;
  p = p * s2 + 2.0 / 7.0;
  p = p * s2 + 2.0 / 5.0;
  p = p * s2 + 2.0 / 3.0;
  const double R = s2 * p;
  const double hfsq = 0.5 * f * f;
  const double de = (double)e;
  return de * kLn2Hi + (f - (hfsq - (s * (hfsq + R) + de * kLn2Lo)));
}

// x^y for x > 0 as e^(y ln x).
ICM_HD double pow(double x, double y) { return exp(y * log(x)); }

// sin x and cos x for |x| < 2^20: Cody & Waite reduction by pi/2 (three-part
// constant), then the Taylor series of sin to r^17 and cos to r^18 on
// |r| <= pi/4 (truncation below 1e-19), rotated by the quadrant.
ICM_HD void sincos(double x, double* s_out, double* c_out) {
  const double q = rint(x * kTwoOverPi);
  const double r = ((x - q * kPio2_1) - q * kPio2_2) - q * kPio2_3;
  const double r2 = r * r;
  double ps = -1.0 / 355687428096000.0;  // -1/17!
  ps = ps * r2 + 1.0 / 1307674368000.0;
  ps = ps * r2 - 1.0 / 6227020800.0;
  ps = ps * r2 + 1.0 / 39916800.0;
  ps = ps * r2 - 1.0 / 362880.0;
  ps = ps * r2 + 1.0 / 5040.0;
  ps = ps * r2 - 1.0 / 120.0;
  ps = ps * r2 + 1.0 / 6.0;
  const double sr = r - r * r2 * ps;
  double pc = 1.0 / 6402373705728000.0;  // 1/18!
  pc = pc * r2 - 1.0 / 20922789888000.0;
  pc = pc * r2 + 1.0 / 87178291200.0;
  pc = pc * r2 - 1.0 / 479001600.0;
  pc = pc * r2 + 1.0 / 3628800.0;
  pc = pc * r2 - 1.0 / 40320.0;
  pc = pc * r2 + 1.0 / 720.0;
  pc = pc * r2 - 1.0 / 24.0;
  pc = pc * r2 + 0.5;
  const double cr = 1.0 - r2 * pc;
  const int quad = (int)((int64_t)q & 3);
  double sv, cv;
  switch (quad) {
    case 0: sv = sr; cv = cr; break;
    case 1: sv = cr; cv = -sr; break;
    case 2: sv = -sr; cv = -cr; break;
    default: sv = -cr; cv = sr; break;
  }
  *s_out = sv;
  *c_out = cv;
}

ICM_HD double sin(double x) {
  double s, c;
  sincos(x, &s, &c);
  return s;
}

ICM_HD double cos(double x) {
  double s, c;
  sincos(x, &s, &c);
  return c;
}

}  // namespace icm
}  // namespace aster
