// Blackbody band fractions of the visible boxcar bands and the table the
// kernels interpolate (the role of crates/gravlume-render/src/spectral_lut.rs
// and the band list of crates/gravlume-domain/src/radiation.rs:4-8).
//
// With x = c_2 / (lambda T), the fraction of sigma T^4 radiated between
// wavelengths lambda_lo and lambda_hi is
//     F = (15 / pi^4) * integral over [x(lambda_hi), x(lambda_lo)] of t^3 / (e^t - 1) dt.
// The integral is evaluated from two classical series (no quadrature):
//   * upper tail, x >= 1:  P(x) = sum_k e^{-k x} (x^3/k + 3 x^2/k^2 + 6 x/k^3 + 6/k^4)
//     (term-by-term integration of t^3 sum_k e^{-k t});
//   * lower part, x <= 1:  I(x) = sum_n b_n x^{n+3} / (n + 3), with b_n the
//     coefficients of t / (e^t - 1) = sum_n b_n t^n: b_0 = 1, b_1 = -1/2,
//     b_{2m} = (-1)^{m+1} 2 zeta(2m) / (2 pi)^{2m}, odd b_n = 0 for n > 1.
// Bands wholly at x >= 50 are taken in logarithmic form,
// log P(x) = -x + log(x^3 + 3 x^2 + 6 x + 6) (the k = 1 term; the rest is
// below e^-50 relative), so temperatures down to 2^-8 K stay representable
// as log2 fractions.
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/gravlume.h"

namespace gl {

gl_status fail(gl_status st, const std::string& msg);

namespace {

constexpr double kC2 = 0.014387768775039337;  // second radiation constant [m K]
constexpr double kPi = 3.14159265358979323846;
constexpr double kBandNm[3][2] = {{600.0, 700.0}, {500.0, 600.0}, {400.0, 500.0}};
constexpr int kLowTerms = 16;  // b_0 .. b_30, enough for x <= 1 to 1e-20

struct LowSeries {
  double b[2 * kLowTerms + 1];
  LowSeries() {
    std::memset(b, 0, sizeof(b));
    b[0] = 1.0;
    b[1] = -0.5;
    const double two_pi = 2.0 * kPi;
    for (int m = 1; m <= kLowTerms - 1; ++m) {
      double zeta;
      if (m == 1) {
        zeta = kPi * kPi / 6.0;
      } else {
        // direct sum to K plus the Euler-Maclaurin tail K^{1-2m}/(2m-1) - K^{-2m}/2
        const int K = 2000;
        zeta = 0.0;
        for (int k = K; k >= 1; --k) zeta += std::pow(double(k), -2.0 * m);
        zeta += std::pow(double(K), 1.0 - 2.0 * m) / (2.0 * m - 1.0) - 0.5 * std::pow(double(K), -2.0 * m);
      }
      const double v = 2.0 * zeta / std::pow(two_pi, 2.0 * m);
      b[2 * m] = (m % 2 == 1) ? v : -v;
    }
  }
};

const LowSeries& low_series() {
  static const LowSeries s;
  return s;
}

// integral over [0, x] of t^3 / (e^t - 1), 0 <= x <= 1
double lower_part(double x) {
  const LowSeries& s = low_series();
  double sum = 0.0;
  double xp = x * x * x;  // x^{n+3}
  for (int n = 0; n <= 2 * kLowTerms; ++n) {
    sum += s.b[n] * xp / double(n + 3);
    xp *= x;
  }
  return sum;
}

// integral over [x, inf) of t^3 / (e^t - 1), x >= 1
double upper_tail(double x) {
  double sum = 0.0;
  for (int k = 1; k <= 80; ++k) {
    const double kk = double(k);
    const double term = std::exp(-kk * x) *
                        (x * x * x / kk + 3.0 * x * x / (kk * kk) + 6.0 * x / (kk * kk * kk) +
                         6.0 / (kk * kk * kk * kk));
    sum += term;
    if (term <= 1e-20 * sum) break;
  }
  return sum;
}

double log_upper_tail_large(double x) {
  return -x + std::log(((x + 3.0) * x + 6.0) * x + 6.0);
}

// log of the kernel integral over [x0, x1], 0 < x0 < x1
double log_band_integral(double x0, double x1) {
  if (x0 >= 50.0) {
    const double l0 = log_upper_tail_large(x0), l1 = log_upper_tail_large(x1);
    return l0 + std::log1p(-std::exp(l1 - l0));
  }
  double v;
  if (x1 <= 1.0)
    v = lower_part(x1) - lower_part(x0);
  else if (x0 >= 1.0)
    v = upper_tail(x0) - upper_tail(x1);
  else
    v = (lower_part(1.0) - lower_part(x0)) + (upper_tail(1.0) - upper_tail(x1));
  return std::log(v);
}

}  // namespace

// log2 fractions of the three bands; false when not representable.
bool blackbody_log2_fractions(double temperature, double out[3]) {
  if (!(temperature > 0.0) || !std::isfinite(temperature)) return false;
  const double log_total = std::log(kPi * kPi * kPi * kPi / 15.0);
  for (int b = 0; b < 3; ++b) {
    const double x0 = kC2 / (kBandNm[b][1] * 1e-9 * temperature);
    const double x1 = kC2 / (kBandNm[b][0] * 1e-9 * temperature);
    if (!std::isfinite(x0) || !std::isfinite(x1) || !(x0 > 0.0) || !(x1 > x0)) return false;
    const double v = (log_band_integral(x0, x1) - log_total) / std::log(2.0);
    if (!std::isfinite(v) || v > 0.0) return false;
    out[b] = v;
  }
  return true;
}

}  // namespace gl

extern "C" {

gl_status gl_blackbody_log2_fractions(double temperature_kelvin, double* out_log2) {
  if (!out_log2) return gl::fail(GL_ERR_INVALID_ARGUMENT, "gl_blackbody_log2_fractions: null output");
  if (!gl::blackbody_log2_fractions(temperature_kelvin, out_log2))
    return gl::fail(GL_ERR_INVALID_ARGUMENT,
                    "gl_blackbody_log2_fractions: temperature must be finite and positive");
  return GL_OK;
}

gl_status gl_blackbody_lut(float* out, size_t n_floats) {
  if (!out || n_floats < size_t(GL_BLACKBODY_LUT_ENTRIES) * 4)
    return gl::fail(GL_ERR_INVALID_ARGUMENT, "gl_blackbody_lut: the table needs 4 floats per entry");
  for (int i = 0; i < GL_BLACKBODY_LUT_ENTRIES; ++i) {
    double f[3];
    const double t = std::exp2(GL_BLACKBODY_LUT_MIN_LOG2_K + double(i) / GL_BLACKBODY_LUT_PER_OCTAVE);
    if (!gl::blackbody_log2_fractions(t, f))
      return gl::fail(GL_ERR_NOT_REPRESENTABLE, "gl_blackbody_lut: band fraction not representable");
    out[4 * i + 0] = float(f[0]);
    out[4 * i + 1] = float(f[1]);
    out[4 * i + 2] = float(f[2]);
    out[4 * i + 3] = 0.0f;
  }
  return GL_OK;
}

}  // extern "C"
