// Oracle: blackbody bands and homogeneous-slab transport of an observed
// equatorial source, f64 (test infrastructure only; see oracle.hpp).
//
// Restates crates/gravlume-reference/src/radiation.rs (the strict CPU path:
// adaptive Simpson over unit intervals of the dimensionless Planck kernel, a
// closed logarithmic tail for x >= 50) and the composition of
// crates/gravlume-reference/src/surface.rs:12-60, with the band table of
// crates/gravlume-domain/src/radiation.rs:4-8 and the slab of :54-220.
//
// Pinned by tests/test_oracle_radiation.py: the four reference fixtures
// fixtures/v3/kerr-blackbody-*.toml (80-digit expected band intensities and
// transported bolometric intensities, 1e-12 absolute), the reference's 6000 K
// and 220 K known answers (radiation.rs:249-287), the Stefan-Boltzmann
// closure of the kernel integral (pi^4/15 over [0, 80]), and an independent
// series evaluation of the band integrals.
#include <cmath>
#include <limits>

#include "oracle.hpp"

namespace orc {
namespace {

constexpr double kSecondRadiation = 0.014387768775039337;  // c_2 [m K]
constexpr double kPlanckIntegral = M_PI * M_PI * M_PI * M_PI / 15.0;
constexpr double kHighFrequency = 50.0;
constexpr unsigned kIntervals = 80;
// (lower, upper) wavelength in nm of the red, green, blue boxcar bands.
constexpr double kBands[3][2] = {{600.0, 700.0}, {500.0, 600.0}, {400.0, 500.0}};

// radiation.rs:224-232
double planck_kernel(double x) {
  if (x == 0.0) return 0.0;
  if (x > 50.0) {
    const double e = std::exp(-x);
    return x * x * x * e / (1.0 - e);
  }
  return x * x * x / std::expm1(x);
}

// radiation.rs:220-222
double simpson(double l, double r, double fl, double fm, double fr) {
  return (r - l) * (std::fma(4.0, fm, fl) + fr) / 6.0;
}

// radiation.rs:193-218
double adaptive_simpson(double l, double r, double whole, double tol, unsigned depth) {
  const double m = 0.5 * (l + r);
  const double lm = 0.5 * (l + m), rm = 0.5 * (m + r);
  const double li = simpson(l, m, planck_kernel(l), planck_kernel(lm), planck_kernel(m));
  const double ri = simpson(m, r, planck_kernel(m), planck_kernel(rm), planck_kernel(r));
  const double refined = li + ri;
  const double corr = refined - whole;
  if (depth == 0 || std::fabs(corr) <= 15.0 * tol) return refined + corr / 15.0;
  return adaptive_simpson(l, m, li, 0.5 * tol, depth - 1) +
         adaptive_simpson(m, r, ri, 0.5 * tol, depth - 1);
}

// radiation.rs:122-145
double integrate_kernel(double lower, double upper) {
  double sum = 0.0;
  for (unsigned i = 0; i < kIntervals; ++i) {
    const double l = std::fmax(lower, double(i)), r = std::fmin(upper, double(i + 1));
    if (!(l < r)) continue;
    const double m = 0.5 * (l + r);
    const double whole = simpson(l, r, planck_kernel(l), planck_kernel(m), planck_kernel(r));
    const double tol =
        8.0 * std::numeric_limits<double>::epsilon() * std::fmax(std::fabs(whole), 2.2250738585072014e-308);
    sum += adaptive_simpson(l, r, whole, tol, 20);
  }
  return sum;
}

// radiation.rs:153-159
double log_tail(double x) {
  const double inv = 1.0 / x;
  const double poly = inv * std::fma(inv, std::fma(6.0, inv, 6.0), 3.0);
  return std::fma(3.0, std::log(x), -x) + std::log1p(poly);
}

// radiation.rs:147-151
double log_high_frequency_band(double lower, double upper) {
  const double lt = log_tail(lower), ut = log_tail(upper);
  return lt + std::log1p(-std::exp(ut - lt));
}

// radiation.rs:110-120
bool scale_by_log2(double value, double log2_scale, double* out) {
  const double scale = std::exp2(log2_scale);
  const double s = (scale == 0.0 && value > 0.0) ? std::exp2(std::log2(value) + log2_scale)
                                                 : value * scale;
  if (!(std::isfinite(s) && s >= 0.0)) return false;
  *out = s;
  return true;
}

}  // namespace

// radiation.rs:83-101
int band_log2_fraction(double temperature, int band, double* out) {
  const double lo_m = kBands[band][0] * 1.0e-9, hi_m = kBands[band][1] * 1.0e-9;
  const double lower_x = kSecondRadiation / (hi_m * temperature);
  const double upper_x = kSecondRadiation / (lo_m * temperature);
  if (!std::isfinite(lower_x) || !std::isfinite(upper_x) || lower_x < 0.0 || upper_x <= lower_x)
    return 1;
  const double log_integral = lower_x >= kHighFrequency
                                  ? log_high_frequency_band(lower_x, upper_x)
                                  : std::log(integrate_kernel(lower_x, upper_x));
  const double v = (log_integral - std::log(kPlanckIntegral)) / M_LN2;
  if (!(std::isfinite(v) && v <= 0.0)) return 1;
  *out = v;
  return 0;
}

double planck_kernel_integral(double lower, double upper) { return integrate_kernel(lower, upper); }

// radiation.rs:13-30
int blackbody_band_intensities(double intensity, double temperature, double out[3]) {
  if (!std::isfinite(intensity) || intensity < 0.0 || !std::isfinite(temperature) ||
      temperature <= 0.0)
    return 1;
  for (int b = 0; b < 3; ++b) {
    double f;
    if (band_log2_fraction(temperature, b, &f)) return 1;
    if (!scale_by_log2(intensity, f, &out[b])) return 1;
  }
  return 0;
}

// radiation.rs:32-49
int transport_bolometric(double incoming, const Transport& tr, double* out, double* tau,
                         double* transmittance) {
  if (!tr.slab) {
    *out = incoming;
    *tau = 0.0;
    *transmittance = 1.0;
    return 0;
  }
  const double t = std::exp(-tr.optical_depth);
  const double o = std::fma(incoming, t, tr.integrated_emission);
  if (!(std::isfinite(o) && o >= 0.0)) return 1;
  *out = o;
  *tau = tr.optical_depth;
  *transmittance = t;
  return 0;
}

// radiation.rs:51-80
int transport_bands(const double in[3], const Transport& tr, bool blackbody_emitter,
                    double out[3]) {
  if (!blackbody_emitter) return 1;
  if (!tr.slab) {
    for (int b = 0; b < 3; ++b) out[b] = in[b];
    return 0;
  }
  const double t = std::exp(-tr.optical_depth);
  double src[3] = {0.0, 0.0, 0.0};
  if (tr.integrated_emission != 0.0) {
    if (!tr.blackbody) return 1;
    if (blackbody_band_intensities(tr.integrated_emission, tr.temperature, src)) return 1;
  }
  for (int b = 0; b < 3; ++b) {
    out[b] = std::fma(in[b], t, src[b]);
    if (!std::isfinite(out[b]) || out[b] < 0.0) return 1;
  }
  return 0;
}

// Constant source function S over optical depth tau: the slab's integrated
// emission S (1 - e^-tau) (crates/gravlume-domain/src/radiation.rs:196-201).
double constant_source_emission(double optical_depth, double source) {
  return source * -std::expm1(-optical_depth);
}

// The radiance chain of observable_at (surface.rs:12-60) from the source
// radius r (M = 1) and frequency ratio g. Returns a SurfaceError.
int surface_radiance(const Emitter& em, const Transport& tr, double r, double g,
                     SurfaceRadiance* out) {
  const double rr = r / 6.0;
  if (!std::isfinite(rr) || rr <= 0.0) return kSurfEmittedIntensity;
  double ie = em.intensity_at_six;
  for (int i = 0; i < 3; ++i) ie /= rr;
  if (!std::isfinite(ie) || ie < 0.0) return kSurfEmittedIntensity;
  double io = ie;
  for (int i = 0; i < 4; ++i) io *= g;
  if (!std::isfinite(io) || io < 0.0) return kSurfObservedIntensity;
  out->emitted_intensity = ie;
  out->vacuum_intensity = io;
  out->emitted_temperature = out->observed_temperature = 0.0;
  out->bands[0] = out->bands[1] = out->bands[2] = 0.0;
  if (em.blackbody) {
    const double te = em.temperature_at_six / std::sqrt(rr * std::sqrt(rr));
    if (!std::isfinite(te) || te <= 0.0) return kSurfTemperature;
    const double to = te * g;
    if (!std::isfinite(to) || to <= 0.0) return kSurfTemperature;
    out->emitted_temperature = te;
    out->observed_temperature = to;
    double vb[3];
    if (blackbody_band_intensities(io, to, vb)) return kSurfSpectrum;
    if (transport_bands(vb, tr, true, out->bands)) return kSurfSpectrum;
  }
  double t, tau;
  if (transport_bolometric(io, tr, &out->observed_intensity, &tau, &t)) return kSurfObservedIntensity;
  out->optical_depth = tau;
  return kSurfOk;
}

}  // namespace orc
