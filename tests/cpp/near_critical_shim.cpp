// Host shim over gravlume_amd/csrc/near_critical.h for the CPU tests
// (tests/test_near_critical_host.py): the margin of f64 initial states and
// the f32 start error, compiled with g++ and called through ctypes.
#include <cstddef>

#include "../../gravlume_amd/csrc/near_critical.h"

extern "C" int nc_state_margin(const double* states8, size_t n, double spin, double charge2,
                               double chart, double r_plus, double* mu, double* r_m) {
  for (size_t i = 0; i < n; ++i) {
    const gl::RadialMargin<double> m =
        gl::state_margin(states8 + 8 * i, spin, charge2, chart, r_plus);
    if (!m.ok) return int(i) + 1;
    mu[i] = m.mu;
    r_m[i] = m.r_m;
  }
  return 0;
}

extern "C" void nc_radial_margin(double E, double L, double Q, double spin, double charge2,
                                 double r_plus, double r_start, double* out3) {
  const gl::RadialMargin<double> m =
      gl::radial_margin<double>(E, L, Q, spin, charge2, r_plus, r_start);
  out3[0] = m.mu;
  out3[1] = m.r_m;
  out3[2] = m.ok ? 1.0 : 0.0;
}

extern "C" double nc_f32_start_error(const double* state8, double spin, double charge2,
                                     double chart) {
  return gl::f32_start_error(state8, spin, charge2, chart);
}

extern "C" double nc_f32_start_error_bound(const double* state8, double spin, double charge2,
                                           double chart) {
  return gl::f32_start_error_bound(state8, spin, charge2, chart);
}
