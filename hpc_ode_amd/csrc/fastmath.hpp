// FP64 natural logarithm for the kernel's hot paths (ln T of every RHS, the Troe falloff
// logarithms, the step-size control powers), written from the textbook method rather than
// calling the vendor library, whose general-purpose log costs about three times as many
// instructions.
//
// Method: x = m 2^e with m in [sqrt(1/2), sqrt(2)); ln m = 2 atanh(s), s = (m - 1) / (m + 1),
// |s| <= 0.1716, summed as 2 s (1 + z/3 + z^2/5 + ... + z^9/19), z = s^2 (the first omitted
// term is below 2^-55 of the sum); ln x = e ln2 + ln m with ln2 split into a high part whose
// product with any exponent |e| <= 1074 is exact and a low correction. Error: at most 2 ulp
// against the C library's log over 4e7 samples (tests/test_fastmath.py).
// Special values follow the C library: ln 0 = -inf, ln(+inf) = +inf, ln(x < 0) = NaN, and
// NaN propagates.
//
// The same source compiles for the host (tests pin it against the C library) and the device.
#pragma once

#include <cmath>
#include <limits>

#if defined(__HIPCC__)
#define HPO_HD __host__ __device__ inline
#else
#define HPO_HD inline
#endif

namespace hpo {

HPO_HD double fast_log(double x) {
  int e = 0;
  double m = std::frexp(x, &e);  // x = m 2^e, m in [0.5, 1) for finite nonzero x
  const bool lo = m < 0.70710678118654752440;
  m = lo ? m + m : m;
  e = lo ? e - 1 : e;
  const double s = (m - 1.0) / (m + 1.0);  // m - 1 is exact for m in [0.5, 2]
  const double z = s * s;
  double p = 1.0 / 19.0;
  p = std::fma(p, z, 1.0 / 17.0);
  p = std::fma(p, z, 1.0 / 15.0);
  p = std::fma(p, z, 1.0 / 13.0);
  p = std::fma(p, z, 1.0 / 11.0);
  p = std::fma(p, z, 1.0 / 9.0);
  p = std::fma(p, z, 1.0 / 7.0);
  p = std::fma(p, z, 1.0 / 5.0);
  p = std::fma(p, z, 1.0 / 3.0);
  const double s2 = s + s;
  const double lm = std::fma(s2 * z, p, s2);
  const double de = static_cast<double>(e);
  // ln 2 = LN2_HI + LN2_LO; LN2_HI has 21 trailing zero bits, so de * LN2_HI is exact
  constexpr double LN2_HI = 6.93147180369123816490e-01;
  constexpr double LN2_LO = 1.90821492927058770002e-10;
  const double r = std::fma(de, LN2_HI, std::fma(de, LN2_LO, lm));
  const double inf = std::numeric_limits<double>::infinity();
  if (x > 0.0 && x < inf) return r;
  if (x == 0.0) return -inf;
  if (x == inf) return inf;
  return std::numeric_limits<double>::quiet_NaN();
}

HPO_HD double fast_log10(double x) { return fast_log(x) * 0.43429448190325182765; }

// x^y for x >= 0 as exp(y ln x); used for the step-size ratios of the BDF controller, where
// x is a positive error-norm ratio and y = 1/L with L = 2..6.
HPO_HD double fast_pow(double x, double y) { return std::exp(y * fast_log(x)); }

}  // namespace hpo
