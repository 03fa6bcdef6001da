// fw_math.hpp -- FP64 exp and log for the reference-arithmetic path.
//
// The reference's numba kernels call the host libm.  A GPU libm that is
// accurate to about one ulp rounds differently from it on a noticeable
// fraction of arguments, and the TP06 upstroke amplifies those last-bit
// differences.  These two functions follow the table-driven methods of
// P. T. P. Tang (ACM TOMS 15(2), 1989, exp; ACM TOMS 16(4), 1990, log):
// argument reduction against a 128-entry double-double table, a short
// polynomial for the remainder, and double-double accumulation, so the value
// before the final rounding carries about 2^-65 relative error.  The final
// result is therefore correctly rounded except for arguments whose exact
// value lies within ~2^-12 ulp of a rounding boundary.  The same code runs on
// the host (tests) and on the device.  Results in the subnormal range are
// rounded twice (table product, then scaling).
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

FW_HD inline double dd_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893383973096) return HUGE_VAL;
  if (x < -745.13321910194122) return 0.0;
  const double kd = rint(x * mt::kInvLn2x128);
  const int k = (int)kd;
  // r1 exact: kd * kLn2d128Hi is exact (<= 18 + 32 bits) and close to x
  const double r1 = fma(-kd, mt::kLn2d128Hi, x);
  const double r2 = -kd * mt::kLn2d128Lo;
  const double r = r1 + r2;
  const double rlo = (r1 - r) + r2;
  const int j = k & 127;
  const int m = (k - j) / 128;
  // exp(r) - 1 - r
  const double q =
      r * r *
      (0.5 + r * (1.0 / 6 + r * (1.0 / 24 + r * (1.0 / 120 + r * (1.0 / 720)))));
  const double th = mt::kExp2Hi[j], tl = mt::kExp2Lo[j];
  // (th + tl) * (1 + r + rlo + q)
  const double prod = th * r;
  const double perr = fma(th, r, -prod);
  const double s = th + prod;
  const double e = (th - s) + prod;
  const double tail = e + perr + th * (rlo + q) + tl * (1.0 + (r + (rlo + q)));
  const double y = s + tail;
  return ldexp(y, m);
}

FW_HD inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

FW_HD inline double dd_log(double x) {
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
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
  const double u = f / F;
  const double ulo = fma(-u, F, f) / F;
  // log1p(u) - u
  const double q =
      u * u *
      (-0.5 +
       u * (1.0 / 3 +
            u * (-0.25 + u * (0.2 + u * (-1.0 / 6 + u * (1.0 / 7 + u * (-0.125 + u * (1.0 / 9))))))));
  const double mh = (double)m * mt::kLn2Hi;  // exact
  const double ml = (double)m * mt::kLn2Lo;
  double s1, e1, s2, e2;
  two_sum(mh, mt::kLogHi[j - 64], s1, e1);
  two_sum(s1, u, s2, e2);
  const double tail = e1 + e2 + ml + mt::kLogLo[j - 64] + (ulo - u * ulo) + q;
  return s2 + tail;
}

}  // namespace fw
