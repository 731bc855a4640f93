// Oracle: CPU f64 restatement of Gravlume's reference null-geodesic trace.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP trace in
// gravlume_amd/. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only to check or time; the product path
// never calls it.
//
// It follows the reference's CPU oracle (crates/gravlume-reference and the f64
// geometry in crates/gravlume-domain); each function cites the reference
// file:line it restates. Arithmetic order follows the reference where that
// order is observable in the last bits (fused multiply-adds are std::fma where
// the reference writes mul_add).
//
// Pinning (see tests/test_oracle_*.py):
//   * reference fixtures fixtures/v1/*.toml (80-digit Schwarzschild quadratures,
//     Kerr observation frame) and fixtures/v2/kerr-surface-observable.toml;
//   * closed forms: r_+ = M + sqrt(M^2 - a^2 - Q^2), Schwarzschild r=6 emitter
//     g = sqrt(1/2), BL metric pull-back, RN/Minkowski limits, BPT ISCO,
//     weak-field deflection 4M/b;
//   * invariants: E, L_z, Carter conservation, time-origin invariance,
//     forward/backward turning radius symmetry.
// Functions without such a check are marked "parity unpinned" below.
#pragma once

#include <array>
#include <cstdint>

namespace orc {

using V3 = std::array<double, 3>;
using V4 = std::array<double, 4>;
using S8 = std::array<double, 8>;

enum Extremality : int { kSubextremal = 0, kExtremal = 1, kSuperextremal = 2 };

// GeometryError (crates/gravlume-domain/src/spacetime.rs:70-80); 0 = success.
enum GeometryError : int {
  kGeomOk = 0,
  kGeomNonFinite = 1,
  kGeomRingSingularity = 2,
  kGeomChartBoundary = 3,
  kGeomInvalidDenominator = 4,
};

// Termination (crates/gravlume-reference/src/outcome.rs:149-158).
enum Termination : int {
  kTermNone = 0,
  kTermHorizon = 1,
  kTermEscape = 2,
  kTermSingularity = 3,
  kTermStepExhaustion = 4,
  kTermNumericalFailure = 5,
  kTermSurface = 7,
  kTermRejectExhaustion = 8,
};

// NumericalFailure (outcome.rs:129-136); 0 when not a failure.
enum NumericalFailureKind : int {
  kFailNone = 0,
  kFailNonFinite = 1,
  kFailRingSingularity = 2,
  kFailChartBoundary = 3,
  kFailInvalidDenominator = 4,
  kFailMinimumStep = 5,
};

// EventKind bit order (events.rs:9-15, geodesic.rs:616-634).
enum EventBit : uint32_t {
  kEventSingularity = 1u,
  kEventHorizon = 2u,
  kEventSurface = 4u,
  kEventEscape = 8u,
};

// SurfaceObservableError (surface.rs:292-316); 0 = success.
enum SurfaceError : int {
  kSurfOk = 0,
  kSurfGeometry = 1,
  kSurfOutside = 2,
  kSurfOrbitUnavailable = 3,
  kSurfNotTimelike = 4,
  kSurfEmitterFrequency = 5,
  kSurfObserverFrequency = 6,
  kSurfRatio = 7,
  kSurfEmittedIntensity = 8,
  kSurfObservedIntensity = 9,
  kSurfTemperature = 10,
  kSurfSpectrum = 11,
};

struct Spacetime {
  double mass = 1.0, spin = 0.0, charge = 0.0;
  double chart = 1.0;  // +1 ingoing, -1 outgoing (spacetime.rs:31-38)
  int extremality = kSubextremal;
  bool has_horizon = true;
  double outer_horizon = 2.0;
};

// ReferencePolicy (crates/gravlume-reference/src/policy.rs:5-66).
struct Policy {
  double pos_rtol, pos_atol, mom_rtol, mom_atol;
  double initial_step, minimum_step, maximum_step;
  double safety, min_factor, max_factor;
  uint64_t max_accepted;
  uint32_t max_rejects;
  double event_affine_tol, event_tie_tol, arming_band, guard_d_over_m4;
  int id;  // 0 regular-v1, 1 strict-v1, 2 custom
  // Not in the reference policy (0 there): when positive, a step is also
  // capped at this multiple of the oblate radius of the state it starts from
  // (the rejected step's start, or the accepted step's end), the GPU trace's
  // step cap, for a CPU baseline that takes the GPU's step sequence.
  double maximum_step_per_radius;
};

// EventConfiguration (events.rs:17-105) plus the horizon offset used by the
// capture surface r_+ (1 + offset); offset 0 restates the reference exactly.
struct Events {
  int has_escape = 1;
  double escape_radius = 200.0;
  int has_surface = 0;
  double surface_inner = 0.0, surface_outer = 0.0;
  double horizon_offset = 0.0;
};

// Emitter (crates/gravlume-domain/src/emission.rs:12-145).
struct Emitter {
  double intensity_at_six = 1.0;
  int blackbody = 0;
  double temperature_at_six = 0.0;
};

// SurfaceTransport (crates/gravlume-domain/src/scene.rs:40-43) with the
// HomogeneousScalarSlab (radiation.rs:54-220): slab = 0 is vacuum; otherwise
// I_out = I_in exp(-optical_depth) + integrated_emission, whose spectrum is a
// blackbody at `temperature` when blackbody = 1.
struct Transport {
  int slab = 0;
  double optical_depth = 0.0;
  double integrated_emission = 0.0;
  int blackbody = 0;
  double temperature = 0.0;
};

// The radiance fields of SurfaceObservable (surface.rs:235-290).
struct SurfaceRadiance {
  double emitted_intensity, vacuum_intensity, observed_intensity, optical_depth;
  double emitted_temperature, observed_temperature;
  double bands[3];  // red, green, blue (blackbody emitters only)
};

struct Outcome {
  int termination;
  int failure;
  uint32_t candidates;
  int polar_side;  // 0 negative, 1 equatorial, 2 positive
  uint32_t radial_turnings, equatorial_crossings;
  int32_t azimuth_winding;
  int has_turning_radius;
  double turning_radius;
  double azimuth_advance;
  double travel_time;
  double affine;
  double state[8];
  uint64_t accepted, rejected, rhs_evaluations;
  double min_step, max_step;
  double drift_null, drift_energy, drift_lz, drift_carter;
  int has_event;
  double event_affine, event_bracket, event_residual;
  int has_escape_direction;
  double escape_direction[3];
  // Observed surface (ObservationTracer only).
  int surface_status;  // -1 not evaluated, else SurfaceError
  double source_radius, source_azimuth, frequency_ratio;
  double emitted_intensity, observed_intensity;
  double emitted_temperature, observed_temperature;
};

// Validated stationary observer and perspective view
// (crates/gravlume-domain/src/{observer,scene,view}.rs).
struct Observation {
  Spacetime st;
  double event[4];
  double u[4], right[4], up[4], arrival[4];
  double gram_residual, orientation;
  int used_up_fallback;
  double g_tt;
  double metric[4][4];
  double omega;
  uint32_t width, height;
  double fov, tan_half_fov;
};

Policy regular_v1();
Policy strict_v1();

int make_spacetime(double mass, double spin, double charge, double chart, Spacetime* out);
int compare_extremality(double mass, double spin, double charge);

struct Geometry {
  double radius;
  V3 radius_gradient;
  double f;
  V3 f_gradient;
  V4 null_covector, null_vector;
  std::array<V4, 3> null_vector_gradient;
};

int geometry(const Spacetime& st, const double* xyz, Geometry* out);
int hamiltonian_rhs(const Spacetime& st, const S8& state, S8* out);
struct Invariants {
  double energy, lz, carter, null_residual;
};
int invariants(const Spacetime& st, const S8& state, Invariants* out);
int radial_velocity(const Spacetime& st, const S8& state, double* out);
int singularity_measure(const Spacetime& st, const double* xyz, double* out);
void metric_covariant(const Geometry& g, double out[4][4]);
void metric_inverse(const Geometry& g, double out[4][4]);
double normalized_quadratic_form_residual(const double m[4][4], const double v[4]);

void trace(const Spacetime& st, const Policy& pol, const Events& ev, const S8& state,
           double direction, Outcome* out);

int build_observation(double mass, double spin, double charge, double chart,
                      const double event[4], const double target[4], const double up_hint[3],
                      double omega, uint32_t width, uint32_t height, double fov,
                      Observation* out);
// Initial ray for an arbitrary sight-plane point (sx, sy) (scene.rs:211-271).
int initial_ray_sight(const Observation& obs, double sx, double sy, S8* state, double sight[4],
                      double* frequency, double* null_residual);
int sight_plane(const Observation& obs, uint32_t px, uint32_t py, double subx, double suby,
                double* sx, double* sy);

int observable_at(const Spacetime& st, const Emitter& em, double inner, double outer,
                  const S8& state, double observer_frequency, Outcome* out);

double isco_prograde(double spin, double charge);

// radiation.cpp: log2 of the blackbody fraction in band 0/1/2 (red, green,
// blue boxcars) at temperature T; 0 on success.
int band_log2_fraction(double temperature, int band, double* out);
double planck_kernel_integral(double lower, double upper);
int blackbody_band_intensities(double intensity, double temperature, double out[3]);
int transport_bolometric(double incoming, const Transport& tr, double* out, double* tau,
                         double* transmittance);
int transport_bands(const double in[3], const Transport& tr, bool blackbody_emitter,
                    double out[3]);
double constant_source_emission(double optical_depth, double source);
int surface_radiance(const Emitter& em, const Transport& tr, double r, double g,
                     SurfaceRadiance* out);

// Forward-mode tangents (tangent.cpp): g and dg/d(a, Q, inclination, s_x, s_y)
// of the backward ray through sight-plane point (sx, sy) of the observation
// at (a, Q, inclination) with the observer at oblate radius r_obs, azimuth 0,
// looking at the origin with +z up. Returns 0 on a surface hit.
int tangent_frequency_ratio(double a, double q, double incl, double chart, double r_obs,
                            uint32_t width, uint32_t height, double fov, double sx, double sy,
                            double inner, double outer, double escape, double rtol, double* g,
                            double dg[5]);

}  // namespace orc
