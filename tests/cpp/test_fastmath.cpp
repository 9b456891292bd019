// Pins hpo::fast_log / fast_log10 / fast_pow (hpc_ode_amd/csrc/fastmath.hpp, the logarithm the
// BDF kernel uses for ln T, the Troe falloff logarithms and the step-size powers) against the C
// library on the host: prints the largest error in ulps over random positive doubles of every
// exponent (denormals included) and over the reduced range [0.5, 2], and the special values.
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
  std::printf("log_bits %.3f\nlog_reduced %.3f\nlog10 %.3f\npow %.3f\n", w_bits, w_red, w_l10, w_pow);
  std::printf("log0 %g\nloginf %g\nlogneg %g\nlognan %g\npow0 %g\n", hpo::fast_log(0.0), hpo::fast_log(INFINITY),
              hpo::fast_log(-1.0), hpo::fast_log(NAN), hpo::fast_pow(0.0, 0.5));
  return 0;
}
