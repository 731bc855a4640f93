// Oracle (test infrastructure only): C entry points used by the Python test
// harness through ctypes, and the order-preserving thread-pool batch that
// restates GeodesicBatch::trace_ordered (crates/gravlume-reference/src/batch.rs:19-43).
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.hpp"

using namespace orc;

namespace {

constexpr unsigned kMaxThreads = 256;  // batch.rs:7

template <class F>
int parallel_ordered(size_t n, unsigned threads, F&& body) {
  if (threads == 0 || threads > kMaxThreads) return 1;
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    for (;;) {
      size_t i = next.fetch_add(1);
      if (i >= n) return;
      body(i);
    }
  };
  if (threads == 1) {
    worker();
    return 0;
  }
  std::vector<std::thread> pool;
  pool.reserve(threads);
  for (unsigned t = 0; t < threads; ++t) pool.emplace_back(worker);
  for (auto& th : pool) th.join();
  return 0;
}

}  // namespace

extern "C" {

int orc_policy_preset(int which, Policy* out) {
  if (which == 0)
    *out = regular_v1();
  else if (which == 1)
    *out = strict_v1();
  else
    return 1;
  return 0;
}

int orc_make_spacetime(double mass, double spin, double charge, double chart, Spacetime* out) {
  return make_spacetime(mass, spin, charge, chart, out);
}

int orc_geometry(const Spacetime* st, const double* xyz, double* radius, double* f,
                 double* radius_gradient, double* f_gradient, double* null_covector,
                 double* null_vector_gradient) {
  Geometry g;
  int e = geometry(*st, xyz, &g);
  if (e) return e;
  *radius = g.radius;
  *f = g.f;
  for (int i = 0; i < 3; ++i) {
    radius_gradient[i] = g.radius_gradient[i];
    f_gradient[i] = g.f_gradient[i];
    for (int j = 0; j < 4; ++j) null_vector_gradient[4 * i + j] = g.null_vector_gradient[i][j];
  }
  for (int i = 0; i < 4; ++i) null_covector[i] = g.null_covector[i];
  return 0;
}

int orc_metric(const Spacetime* st, const double* xyz, double* cov16, double* inv16) {
  Geometry g;
  int e = geometry(*st, xyz, &g);
  if (e) return e;
  double c[4][4], iv[4][4];
  metric_covariant(g, c);
  metric_inverse(g, iv);
  std::memcpy(cov16, c, sizeof c);
  std::memcpy(inv16, iv, sizeof iv);
  return 0;
}

int orc_rhs(const Spacetime* st, const double* state8, double* out8) {
  S8 y, d;
  std::memcpy(y.data(), state8, sizeof(double) * 8);
  int e = hamiltonian_rhs(*st, y, &d);
  if (e) return e;
  std::memcpy(out8, d.data(), sizeof(double) * 8);
  return 0;
}

int orc_invariants(const Spacetime* st, const double* state8, double* out4) {
  S8 y;
  std::memcpy(y.data(), state8, sizeof(double) * 8);
  Invariants inv;
  int e = invariants(*st, y, &inv);
  if (e) return e;
  out4[0] = inv.energy;
  out4[1] = inv.lz;
  out4[2] = inv.carter;
  out4[3] = inv.null_residual;
  return 0;
}

int orc_radial_velocity(const Spacetime* st, const double* state8, double* out) {
  S8 y;
  std::memcpy(y.data(), state8, sizeof(double) * 8);
  return radial_velocity(*st, y, out);
}

void orc_trace(const Spacetime* st, const Policy* pol, const Events* ev, const double* state8,
               double direction, Outcome* out) {
  S8 y;
  std::memcpy(y.data(), state8, sizeof(double) * 8);
  trace(*st, *pol, *ev, y, direction, out);
}

// GeodesicBatch::trace_ordered: outcome i belongs to request i.
int orc_trace_ordered(const Spacetime* st, const Policy* pol, const Events* ev,
                      const double* states8, const double* directions, size_t n,
                      unsigned threads, Outcome* out) {
  return parallel_ordered(n, threads, [&](size_t i) {
    S8 y;
    std::memcpy(y.data(), states8 + 8 * i, sizeof(double) * 8);
    trace(*st, *pol, *ev, y, directions[i], &out[i]);
  });
}

int orc_build_observation(double mass, double spin, double charge, double chart,
                          const double* event4, const double* target4, const double* up3,
                          double omega, uint32_t width, uint32_t height, double fov,
                          Observation* out) {
  return build_observation(mass, spin, charge, chart, event4, target4, up3, omega, width, height,
                           fov, out);
}

int orc_sight_plane(const Observation* obs, uint32_t px, uint32_t py, double subx, double suby,
                    double* sxy) {
  return sight_plane(*obs, px, py, subx, suby, &sxy[0], &sxy[1]);
}

int orc_initial_ray(const Observation* obs, double sx, double sy, double* state8, double* sight4,
                    double* freq_and_null) {
  S8 y;
  int e = initial_ray_sight(*obs, sx, sy, &y, sight4, &freq_and_null[0], &freq_and_null[1]);
  if (e) return e;
  std::memcpy(state8, y.data(), sizeof(double) * 8);
  return 0;
}

int orc_observable_at(const Spacetime* st, const Emitter* em, double inner, double outer,
                      const double* state8, double omega, Outcome* out) {
  S8 y;
  std::memcpy(y.data(), state8, sizeof(double) * 8);
  return observable_at(*st, *em, inner, outer, y, omega, out);
}

// ObservationTracer::trace (crates/gravlume-reference/src/observation.rs:60-102)
// over a batch of sight-plane points, with an optional equatorial emitter.
// The observer frequency must be 1 within 32 ulp, as in the reference.
int orc_observation_trace_ordered(const Observation* obs, const Policy* pol, const Events* ev,
                                  const Emitter* em, const double* sight_xy, size_t n,
                                  unsigned threads, Outcome* out, int* status) {
  return parallel_ordered(n, threads, [&](size_t i) {
    S8 y;
    double sight[4], fn[2];
    int e = initial_ray_sight(*obs, sight_xy[2 * i], sight_xy[2 * i + 1], &y, sight, &fn[0],
                              &fn[1]);
    if (e) {
      status[i] = 10 + e;
      std::memset(&out[i], 0, sizeof(Outcome));
      return;
    }
    if (std::fabs(fn[0] - 1.0) > 32.0 * 2.220446049250313e-16) {
      status[i] = 20;
      std::memset(&out[i], 0, sizeof(Outcome));
      return;
    }
    trace(obs->st, *pol, *ev, y, -1.0, &out[i]);
    status[i] = 0;
    if (ev->has_surface && out[i].termination == kTermSurface && em) {
      S8 ys;
      std::memcpy(ys.data(), out[i].state, sizeof(double) * 8);
      out[i].surface_status =
          observable_at(obs->st, *em, ev->surface_inner, ev->surface_outer, ys, fn[0], &out[i]);
    }
  });
}

// Forward-mode tangents of g for a batch of sight-plane points (tangent.cpp).
int orc_tangent_g_ordered(double a, double q, double incl, double chart, double r_obs,
                          uint32_t width, uint32_t height, double fov, const double* sight_xy,
                          size_t n, double inner, double outer, double escape, double rtol,
                          unsigned threads, double* g, double* dg, int* status) {
  return parallel_ordered(n, threads, [&](size_t i) {
    status[i] = tangent_frequency_ratio(a, q, incl, chart, r_obs, width, height, fov,
                                        sight_xy[2 * i], sight_xy[2 * i + 1], inner, outer, escape,
                                        rtol, &g[i], &dg[5 * i]);
  });
}

double orc_isco_prograde(double spin, double charge) { return isco_prograde(spin, charge); }

// Blackbody band fractions and surface transport (radiation.cpp).
int orc_band_log2_fractions(double temperature, double* out3) {
  for (int b = 0; b < 3; ++b)
    if (band_log2_fraction(temperature, b, &out3[b])) return 1;
  return 0;
}
double orc_planck_kernel_integral(double lower, double upper) {
  return planck_kernel_integral(lower, upper);
}
int orc_blackbody_band_intensities(double intensity, double temperature, double* out3) {
  return blackbody_band_intensities(intensity, temperature, out3);
}
int orc_transport_bolometric(double incoming, const Transport* tr, double* out3) {
  return transport_bolometric(incoming, *tr, &out3[0], &out3[1], &out3[2]);
}
double orc_constant_source_emission(double optical_depth, double source) {
  return constant_source_emission(optical_depth, source);
}
// surface_radiance for n (r, g) pairs; out holds 9 doubles per pair:
// emitted, vacuum, observed intensity, optical depth, emitted and observed
// temperature, three bands.
int orc_surface_radiance(const Emitter* em, const Transport* tr, const double* r, const double* g,
                         size_t n, double* out, int* status) {
  for (size_t i = 0; i < n; ++i) {
    SurfaceRadiance s{};
    status[i] = surface_radiance(*em, *tr, r[i], g[i], &s);
    const double v[9] = {s.emitted_intensity, s.vacuum_intensity, s.observed_intensity,
                         s.optical_depth, s.emitted_temperature, s.observed_temperature,
                         s.bands[0], s.bands[1], s.bands[2]};
    for (int k = 0; k < 9; ++k) out[9 * i + k] = status[i] ? 0.0 : v[k];
  }
  return 0;
}
size_t orc_sizeof_transport() { return sizeof(Transport); }
size_t orc_sizeof_emitter() { return sizeof(Emitter); }

int orc_compare_extremality(double m, double a, double q) { return compare_extremality(m, a, q); }

size_t orc_sizeof_outcome() { return sizeof(Outcome); }
size_t orc_sizeof_observation() { return sizeof(Observation); }
size_t orc_sizeof_policy() { return sizeof(Policy); }
size_t orc_sizeof_spacetime() { return sizeof(Spacetime); }
size_t orc_sizeof_events() { return sizeof(Events); }

}  // extern "C"

extern "C" {
// KerrNewmanSpacetime::oblate_to_cartesian (crates/gravlume-domain/src/spacetime.rs:245-254).
void orc_oblate_to_cartesian(double spin, double chart, double radius, double polar,
                             double azimuth, double* xyz) {
  double st = std::sin(polar), ct = std::cos(polar);
  double sp = std::sin(azimuth), cp = std::cos(azimuth);
  double cs = (chart >= 0.0 ? 1.0 : -1.0) * spin;
  xyz[0] = std::fma(cs, -sp, radius * cp) * st;
  xyz[1] = std::fma(cs, cp, radius * sp) * st;
  xyz[2] = radius * ct;
}
}
