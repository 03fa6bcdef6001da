// fw_math.hpp -- FP64 exp and log for the reference-arithmetic path.
//
// The reference's numba kernels call the host libm (glibc on Linux), and the
// TP06 upstroke amplifies last-bit differences, so these functions follow the
// host libm's results rather than the GPU library's.
//
// exp: the table-driven method of Szabolcs Nagy's exp in Arm Optimized
// Routines (2018), which glibc adopted in 2.28 (sysdeps/ieee754/dbl-64/e_exp.c):
// k = round(x * 128 / ln 2), r = x - k ln 2 / 128 with a two-part constant,
// 2^(k/128) from a 128-entry table as H (1 + T), exp(r) - 1 by a degree-5
// polynomial with the method's published minimax coefficients, and the result
// as scale + scale * tmp.  Multiply-adds are fused where an x86-64 build with
// FMA (the code glibc selects on any CPU that has FMA) fuses them; with that
// the results equal the host libm's bit for bit (checked on 2e7 random bit
// patterns and every special case; the non-FMA build differs in ~7e-4 of
// arguments).
//
// log: the table-driven method of P. T. P. Tang (ACM TOMS 16(4), 1990):
// reduction against a 128-entry double-double table, a short polynomial and
// double-double accumulation, so the value before the final rounding carries
// about 2^-65 relative error; correctly rounded except within ~2^-12 ulp of a
// rounding boundary.
//
// The same code runs on the host (tests) and on the device.
#pragma once

#include <math.h>
#include <stdint.h>

#include "fw_math_tables.hpp"

#if defined(__HIPCC__)
#define FW_HD __host__ __device__
#else
#define FW_HD
#endif

namespace fw {

namespace mt {
constexpr double kShift = 0x1.8p52;  // round-to-integer shifter
// minimax coefficients of exp(r) - 1 - r on |r| <= ln 2 / 256, as published
constexpr double kC2 = 0x1.ffffffffffdbdp-2;
constexpr double kC3 = 0x1.555555555543cp-3;
constexpr double kC4 = 0x1.55555cf172b91p-5;
constexpr double kC5 = 0x1.1111167a4d017p-7;
}  // namespace mt

FW_HD inline uint64_t f64_bits(double x) { return __builtin_bit_cast(uint64_t, x); }
FW_HD inline double f64_from(uint64_t u) { return __builtin_bit_cast(double, u); }

// exp(x): the reduction and polynomial below run for every argument; a
// wavefront branches to the fix-up of t128_exp_special only when one of its
// lanes has |x| >= 512 (or inf, NaN): one compare of |x| per call.  The fix-up
// is inlined: a call would make the caller keep its live registers across
// 70-odd call sites.  Tiny arguments need no fix-up: for |x| < 2^-54 the
// reduction gives k = 0 and r = x, the polynomial's terms beyond r vanish
// against x, and the result is fma(1, x, 1) = RN(1 + x), the fix-up's value.
FW_HD inline double t128_exp_special(double x, double tmp, uint64_t sbits, uint64_t ki) {
  const uint64_t ux = f64_bits(x);
  const uint32_t abstop = (uint32_t)(ux >> 52) & 0x7ffu;
  if (abstop < 0x3c9u) return 1.0 + x;  // |x| < 2^-54
  if (abstop >= 0x409u) {               // |x| >= 1024, inf, NaN
    if (ux == f64_bits(-HUGE_VAL)) return 0.0;
    if (abstop >= 0x7ffu) return 1.0 + x;
    return (ux >> 63) ? 0.0 : HUGE_VAL;
  }
  // 512 <= |x| < 1024: the scale's exponent may leave the normal range
  if ((ki & 0x80000000u) == 0) {
    const double scale = f64_from(sbits - (1009ull << 52));
    return 0x1p1009 * fma(scale, tmp, scale);
  }
  const double scale = f64_from(sbits + (1022ull << 52));
  const double st = scale * tmp;  // not fused here: two uses
  double y = scale + st;
  if (y < 1.0) {
    // round to the subnormal precision before scaling (no double rounding)
    double lo = scale - y + st;
    const double hi = 1.0 + y;
    lo = 1.0 - hi + y + lo;
    y = (hi + lo) - 1.0;
    if (y == 0.0) y = 0.0;
  }
  return 0x1p-1022 * y;
}

FW_HD inline double t128_exp(double x) {
  double kd = fma(mt::kInvLn2N, x, mt::kShift);
  const uint64_t ki = f64_bits(kd);
  kd -= mt::kShift;
  const double r = fma(kd, mt::kNegLn2loN, fma(kd, mt::kNegLn2hiN, x));
  const uint32_t j = (uint32_t)ki & 127u;
  const double tail = f64_from(mt::kExpTab[2 * j]);
  const uint64_t sbits = mt::kExpTab[2 * j + 1] + (ki << 45);
  const double r2 = r * r;
  const double tmp =
      fma(r2 * r2, fma(r, mt::kC5, mt::kC4), fma(r2, fma(r, mt::kC3, mt::kC2), tail + r));
  if (__builtin_expect(!(__builtin_fabs(x) < 512.0), 0))  // |x| >= 512, inf, NaN
    return t128_exp_special(x, tmp, sbits, ki);
  const double scale = f64_from(sbits);
  return fma(scale, tmp, scale);
}

FW_HD inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

FW_HD inline double dd_log(double x) {
  if (!(x > 0.0 && x < HUGE_VAL)) {  // zero, negative, infinite, NaN
    if (x == 0.0) return -HUGE_VAL;
    if (x == HUGE_VAL) return x;
    return NAN;
  }
  int ex;
  double y = frexp(x, &ex);  // x = y * 2^ex, y in [0.5, 1)
  y *= 2.0;
  int m = ex - 1;
  if (y > 1.4142135623730951) {
    y *= 0.5;
    m += 1;
  }
  const int j = (int)rint(y * 128.0);  // 90..181
  const double F = (double)j * (1.0 / 128.0);
  const double f = y - F;  // exact (Sterbenz)
  // u + ulo = f / F to about 2^-105 relative: u = f RN(128 / j) is within
  // 1.5 ulp of f / F, and the remainder f - u F (one fma) times RN(128 / j)
  // supplies the rest
  const double inv = mt::kLogInvF[j - 64];
  const double u = f * inv;
  const double ulo = fma(-u, F, f) * inv;
  // log1p(u) - u, |u| <= 2^-7.5: its rounding enters at 2^-68 of the result
  const double p = fma(
      u,
      fma(u,
          fma(u, fma(u, fma(u, fma(u, fma(u, 1.0 / 9, -0.125), 1.0 / 7), -1.0 / 6), 0.2),
              -0.25),
          1.0 / 3),
      -0.5);
  const double q = u * u * p;
  const double mh = (double)m * mt::kLn2Hi;  // exact
  const double ml = (double)m * mt::kLn2Lo;
  double s1, e1, s2, e2;
  two_sum(mh, mt::kLogHi[j - 64], s1, e1);
  two_sum(s1, u, s2, e2);
  const double tail = e1 + e2 + ml + mt::kLogLo[j - 64] + (ulo - u * ulo) + q;
  return s2 + tail;
}

}  // namespace fw
