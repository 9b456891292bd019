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

// FP64 exponential by table lookup and a short polynomial (P. T. P. Tang, "Table-driven
// implementation of the exponential function in IEEE floating-point arithmetic", ACM TOMS
// 15(2), 1989): x = (32 m + j) ln2/32 + r with |r| <= ln2/64, exp(x) = 2^m 2^(j/32) exp(r),
// exp(r) - 1 = r + r^2 (1/2 + r/6 + r^2/24 + r^3/120 + r^4/720) (Taylor; the first omitted
// term, r^7/5040, is below 2^-60 of the result). ln2/32 is split Cody-Waite style: LN2_32_HI
// has 32 significant bits, so k LN2_32_HI is exact for every |k| < 2^21, which covers the
// finite range. The table holds 2^(j/32) correctly rounded (kExp2Tab32; on the device a copy
// in LDS is passed in). Error: at most 2 ulp against the C library (tests/test_fastmath.py).
// Results overflow to +inf above 709.78 and underflow through the subnormals to 0 below
// -745.13, as the C library's do; NaN propagates. The reduction holds for |x| up to about
// 1e15 (beyond that x * 32/ln2 no longer rounds to within one of the nearest integer and r
// grows without bound): the kernels' arguments stay within 1e13, the Troe coefficients being
// bounded where the tables are built. Infinite arguments give NaN.
constexpr double kExp2Tab32[32] = {
    0x1.0000000000000p+0, 0x1.059b0d3158574p+0, 0x1.0b5586cf9890fp+0, 0x1.11301d0125b51p+0,
    0x1.172b83c7d517bp+0, 0x1.1d4873168b9aap+0, 0x1.2387a6e756238p+0, 0x1.29e9df51fdee1p+0,
    0x1.306fe0a31b715p+0, 0x1.371a7373aa9cbp+0, 0x1.3dea64c123422p+0, 0x1.44e086061892dp+0,
    0x1.4bfdad5362a27p+0, 0x1.5342b569d4f82p+0, 0x1.5ab07dd485429p+0, 0x1.6247eb03a5585p+0,
    0x1.6a09e667f3bcdp+0, 0x1.71f75e8ec5f74p+0, 0x1.7a11473eb0187p+0, 0x1.82589994cce13p+0,
    0x1.8ace5422aa0dbp+0, 0x1.93737b0cdc5e5p+0, 0x1.9c49182a3f090p+0, 0x1.a5503b23e255dp+0,
    0x1.ae89f995ad3adp+0, 0x1.b7f76f2fb5e47p+0, 0x1.c199bdd85529cp+0, 0x1.cb720dcef9069p+0,
    0x1.d5818dcfba487p+0, 0x1.dfc97337b9b5fp+0, 0x1.ea4afa2a490dap+0, 0x1.f50765b6e4540p+0};

HPO_HD double tab_exp(double x, const double* tab) {
  constexpr double INV_LN2_32 = 0x1.71547652b82fep+5;   // 32 / ln 2
  constexpr double LN2_32_HI = 0x1.62e42ff000000p-6;
  constexpr double LN2_32_LO = -0x1.718432a1b0e26p-40;
  const double k = std::rint(x * INV_LN2_32);
  const int ki = static_cast<int>(fmax(fmin(k, 2147483520.0), -2147483520.0));
  double r = std::fma(-k, LN2_32_HI, x);
  r = std::fma(-k, LN2_32_LO, r);
  double p = 0x1.6c16c16c16c17p-10;           // 1/720
  p = std::fma(p, r, 0x1.1111111111111p-7);   // 1/120
  p = std::fma(p, r, 0x1.5555555555555p-5);   // 1/24
  p = std::fma(p, r, 0x1.5555555555555p-3);   // 1/6
  p = std::fma(p, r, 0.5);
  const double q = std::fma(r * r, p, r);     // exp(r) - 1
  const double t = tab[ki & 31];
  return std::ldexp(std::fma(t, q, t), ki >> 5);
}

// x^y for x >= 0, y > 0 as exp(y ln x); used for the step-size ratios of the BDF controller,
// where x is an error-norm ratio and y = 1/L with L = 2..6. x = 0 (an error estimate of exactly
// zero, e.g. a cell with no reaction to run) gives 0, as pow does: ln 0 = -inf is outside
// tab_exp's domain.
HPO_HD double fast_pow(double x, double y) { return std::exp(y * fast_log(x)); }
HPO_HD double fast_pow(double x, double y, const double* exptab) {
  return x > 0.0 ? tab_exp(y * fast_log(x), exptab) : 0.0;
}

}  // namespace hpo
