// Host build of the product's FP64 exp/log (finitewave_amd/csrc/fw_math.hpp)
// for the CPU tests: each function evaluates one of them beside the host
// libm over an array and counts the results that differ in any bit.
#include <math.h>
#include <stdint.h>

#include "fw_math.hpp"

extern "C" {

int64_t fw_check_exp(const double* x, int64_t n, double* mine, double* libm) {
  int64_t bad = 0;
  for (int64_t i = 0; i < n; ++i) {
    mine[i] = fw::t128_exp(x[i]);
    libm[i] = exp(x[i]);
    const bool both_nan = mine[i] != mine[i] && libm[i] != libm[i];
    if (!both_nan && fw::f64_bits(mine[i]) != fw::f64_bits(libm[i])) ++bad;
  }
  return bad;
}

int64_t fw_check_log(const double* x, int64_t n, double* mine, double* libm) {
  int64_t bad = 0;
  for (int64_t i = 0; i < n; ++i) {
    mine[i] = fw::dd_log(x[i]);
    libm[i] = log(x[i]);
    const bool both_nan = mine[i] != mine[i] && libm[i] != libm[i];
    if (!both_nan && fw::f64_bits(mine[i]) != fw::f64_bits(libm[i])) ++bad;
  }
  return bad;
}

}  // extern "C"
