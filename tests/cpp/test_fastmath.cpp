// Pins hpo::fast_log / fast_log10 / fast_pow and hpo::tab_exp (hpc_ode_amd/csrc/fastmath.hpp,
// the logarithm and exponential the BDF kernel uses for ln T, the rate constants, the
// equilibrium constants, the Troe falloff and the step-size powers) against the C library on
// the host: prints the largest error in ulps over random positive doubles of every exponent
// (denormals included) and over the reduced ranges, and the special values.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "fastmath.hpp"

static double ulps(double a, double ref) {
  if (a == ref) return 0.0;
  const double r = std::fabs(ref);
  return std::fabs(a - ref) / (std::nextafter(r, INFINITY) - r);
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 2000000;
  std::mt19937_64 g(12345);
  double w_bits = 0.0, w_red = 0.0, w_l10 = 0.0, w_pow = 0.0;
  for (long i = 0; i < n; ++i) {
    uint64_t b = g() & 0x7fefffffffffffffull;  // positive, finite, any exponent
    double x;
    std::memcpy(&x, &b, 8);
    if (x == 0.0) continue;
    w_bits = std::fmax(w_bits, ulps(hpo::fast_log(x), std::log(x)));
    w_l10 = std::fmax(w_l10, ulps(hpo::fast_log10(x), std::log10(x)));
  }
  std::uniform_real_distribution<double> red(0.5, 2.0), base(1e-6, 1e3);
  std::uniform_int_distribution<int> L(2, 6);
  for (long i = 0; i < n; ++i) {
    const double x = red(g);
    w_red = std::fmax(w_red, ulps(hpo::fast_log(x), std::log(x)));
    const double bx = base(g), y = 1.0 / L(g);
    w_pow = std::fmax(w_pow, ulps(hpo::fast_pow(bx, y), std::pow(bx, y)));
  }
  // exp over its whole finite-result range (subnormal results measured in units of the
  // smallest subnormal), over [-1, 1], and at every table breakpoint's neighbourhood
  double w_exp = 0.0, w_exp_small = 0.0, w_exp_sub = 0.0, w_tpow = 0.0;
  std::uniform_real_distribution<double> full(-708.0, 709.7), small(-1.0, 1.0), sub(-745.0, -708.5);
  for (long i = 0; i < n; ++i) {
    double x = full(g);
    w_exp = std::fmax(w_exp, ulps(hpo::tab_exp(x, hpo::kExp2Tab32), std::exp(x)));
    x = small(g);
    w_exp_small = std::fmax(w_exp_small, ulps(hpo::tab_exp(x, hpo::kExp2Tab32), std::exp(x)));
    x = sub(g);
    w_exp_sub = std::fmax(w_exp_sub, std::fabs(hpo::tab_exp(x, hpo::kExp2Tab32) - std::exp(x)) / 4.9406564584124654e-324);
    const double bx = base(g), y = 1.0 / L(g);
    w_tpow = std::fmax(w_tpow, ulps(hpo::fast_pow(bx, y, hpo::kExp2Tab32), std::pow(bx, y)));
  }
  for (int m = -1000; m <= 1000; ++m)
    for (int d = -2; d <= 2; ++d) {
      const double x = std::nextafter(m * 0.6931471805599453 / 32.0, d < 0 ? -INFINITY : INFINITY) + d * 1e-17;
      if (x > -708.0 && x < 709.7) w_exp = std::fmax(w_exp, ulps(hpo::tab_exp(x, hpo::kExp2Tab32), std::exp(x)));
    }
  std::printf("log_bits %.3f\nlog_reduced %.3f\nlog10 %.3f\npow %.3f\n", w_bits, w_red, w_l10, w_pow);
  std::printf("exp %.3f\nexp_small %.3f\nexp_subnormal %.3f\ntab_pow %.3f\n", w_exp, w_exp_small, w_exp_sub, w_tpow);
  std::printf("exp0 %g\nexp_big %g\nexp_tiny %g\nexp_nan %g\nexp_max %g\n", hpo::tab_exp(0.0, hpo::kExp2Tab32),
              hpo::tab_exp(800.0, hpo::kExp2Tab32), hpo::tab_exp(-800.0, hpo::kExp2Tab32),
              hpo::tab_exp(NAN, hpo::kExp2Tab32), hpo::tab_exp(709.78, hpo::kExp2Tab32));
  std::printf("exp_neg_huge %g\nexp_pos_huge %g\nexp_neginf %g\nexp_posinf %g\n",
              hpo::tab_exp(-1e13, hpo::kExp2Tab32), hpo::tab_exp(1e13, hpo::kExp2Tab32),
              hpo::tab_exp(-INFINITY, hpo::kExp2Tab32), hpo::tab_exp(INFINITY, hpo::kExp2Tab32));
  std::printf("tab_pow0 %g\n", hpo::fast_pow(0.0, 0.5, hpo::kExp2Tab32));
  std::printf("log0 %g\nloginf %g\nlogneg %g\nlognan %g\npow0 %g\n", hpo::fast_log(0.0), hpo::fast_log(INFINITY),
              hpo::fast_log(-1.0), hpo::fast_log(NAN), hpo::fast_pow(0.0, 0.5));
  return 0;
}
