// Host-side scene validation and f64 -> f32 packing of the device scene
// record: the role of TraceUniforms::from_observation_with_surface
// (crates/gravlume-render/src/trace/input.rs:45-131) together with the
// stationary-observer tetrad (crates/gravlume-domain/src/observer.rs:67-183).
#include "gl_scene.h"

#include <cmath>
#include <cstring>
#include <string>

#include "ks_math.h"

namespace gl {
namespace {

struct V4 {
  double v[4];
};

// Kerr-Schild metric g = eta + f l (x) l at a point, l_mu = (1, l_spatial).
struct Metric {
  double g[4][4];
  double f;
  double l[4];
};

bool metric_at(const gl_scene_desc& d, const double xyz[3], Metric* m) {
  Geom<double> g = geometry<double>(xyz[0], xyz[1], xyz[2], d.mass, d.spin, d.spin * d.spin,
                                    d.charge * d.charge, double(d.chart >= 0 ? 1 : -1));
  if (!g.ok || !std::isfinite(g.f)) return false;
  m->f = g.f;
  m->l[0] = 1.0;
  m->l[1] = g.lx;
  m->l[2] = g.ly;
  m->l[3] = g.lz;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      double eta = a == b ? (a == 0 ? -1.0 : 1.0) : 0.0;
      m->g[a][b] = eta + g.f * m->l[a] * m->l[b];
    }
  return true;
}

double dot(const Metric& m, const V4& a, const V4& b) {
  double s = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) s += m.g[i][j] * a.v[i] * b.v[j];
  return s;
}

// The contraction and the sum of the magnitudes of its sixteen terms
// (metric_dot_and_term_norm, spacetime.rs:463-483): the scale against which
// the Gram check measures a residual.
double dot_term_norm(const Metric& m, const V4& a, const V4& b, double* norm) {
  double s = 0.0, n = 0.0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      const double t = m.g[i][j] * a.v[i] * b.v[j];
      s += t;
      n += std::fabs(t);
    }
  *norm = n;
  return s;
}

// Gram-Schmidt against u (timelike, u.u = -1) and the given spacelike unit
// vectors; false when the remainder is degenerate.
bool orthonormalize(const Metric& m, const V4& u, const V4* basis, int nb, V4 seed, V4* out) {
  double ud = dot(m, u, seed);
  for (int i = 0; i < 4; ++i) seed.v[i] += ud * u.v[i];
  for (int k = 0; k < nb; ++k) {
    double c = dot(m, seed, basis[k]);
    for (int i = 0; i < 4; ++i) seed.v[i] -= c * basis[k].v[i];
  }
  double n2 = dot(m, seed, seed);
  if (!(n2 > 1e-24) || !std::isfinite(n2)) return false;
  double inv = 1.0 / std::sqrt(n2);
  for (int i = 0; i < 4; ++i) out->v[i] = seed.v[i] * inv;
  return true;
}

double det4(const V4 c[4]) {
  double a[4][4];
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k) a[r][k] = c[k].v[r];
  double det = 1.0;
  for (int p = 0; p < 4; ++p) {
    int piv = p;
    for (int r = p + 1; r < 4; ++r)
      if (std::fabs(a[r][p]) > std::fabs(a[piv][p])) piv = r;
    if (a[piv][p] == 0.0) return 0.0;
    if (piv != p) {
      for (int k = 0; k < 4; ++k) std::swap(a[piv][k], a[p][k]);
      det = -det;
    }
    det *= a[p][p];
    for (int r = p + 1; r < 4; ++r) {
      double f = a[r][p] / a[p][p];
      for (int k = p; k < 4; ++k) a[r][k] -= f * a[p][k];
    }
  }
  return det;
}

// Exact comparison of M^2 against a^2 + Q^2 after f32 rounding of (a, Q):
// both sides are formed in long double from exactly representable squares of
// f32 values, so the sign is exact.
int extremality_f32(float mass, float spin, float charge) {
  long double lhs = (long double)mass * mass;
  long double rhs = (long double)spin * spin + (long double)charge * charge;
  return lhs > rhs ? 0 : (lhs == rhs ? 1 : 2);
}

bool set_err(std::string* err, const char* msg) {
  if (err) *err = msg;
  return false;
}

}  // namespace

int extremality_f64(double mass, double spin, double charge) {
  // Two-product splitting makes each square exact; the sum of the six terms
  // is taken in long double after ordering by magnitude, which is exact for
  // the moderate exponents of physical scenes (|value| in [2^-100, 2^100]).
  auto split = [](double v, double* hi, double* lo) {
    *hi = v * v;
    *lo = std::fma(v, v, -*hi);
  };
  double mh, ml, ah, al, qh, ql;
  split(mass, &mh, &ml);
  split(spin, &ah, &al);
  split(charge, &qh, &ql);
  long double d = ((long double)mh - (long double)ah - (long double)qh) +
                  ((long double)ml - (long double)al - (long double)ql);
  return d > 0 ? 0 : (d == 0 ? 1 : 2);
}

namespace {

bool f32_normal_or_zero(double v) {
  const float f = (float)v;
  return std::isfinite(f) && (f == 0.0f || std::fabs(f) >= 1.17549435e-38f) && !(v > 0.0 && f == 0.0f);
}

bool in_lut(double t) {
  return t >= std::exp2(GL_BLACKBODY_LUT_MIN_LOG2_K) &&
         t <= std::exp2(GL_BLACKBODY_LUT_MIN_LOG2_K + (GL_BLACKBODY_LUT_ENTRIES - 1) / GL_BLACKBODY_LUT_PER_OCTAVE);
}

}  // namespace

// The emitter and transport checks of the reference domain and GPU packing:
// emission.rs:71-145 (radii, non-negative intensity, positive temperature),
// radiation.rs:160-220 (non-negative depth and emission, positive
// temperature), scene.rs:58-77 (a blackbody emitter behind an emitting slab
// needs the slab's spectrum), input.rs:166-320 (temperatures inside the band
// table at both radii, values that stay normal in f32).
int check_surface_radiance(const gl_scene_desc& d, std::string* err) {
  if (d.emission_model != GL_EMISSION_BOLOMETRIC && d.emission_model != GL_EMISSION_BLACKBODY)
    return set_err(err, "unknown emission model"), GL_ERR_INVALID_ARGUMENT;
  if (d.transport_model != GL_TRANSPORT_VACUUM && d.transport_model != GL_TRANSPORT_SLAB)
    return set_err(err, "unknown transport model"), GL_ERR_INVALID_ARGUMENT;
  if (!std::isfinite(d.intensity_at_six) || d.intensity_at_six < 0.0)
    return set_err(err, "bolometric intensity must be finite and non-negative"), GL_ERR_INVALID_ARGUMENT;
  if (!f32_normal_or_zero(d.intensity_at_six))
    return set_err(err, "surface intensity not representable in f32"), GL_ERR_NOT_REPRESENTABLE;
  const bool blackbody = d.emission_model == GL_EMISSION_BLACKBODY;
  if (blackbody) {
    if (!std::isfinite(d.temperature_at_six) || !(d.temperature_at_six > 0.0))
      return set_err(err, "blackbody temperature must be finite and positive"), GL_ERR_INVALID_ARGUMENT;
    for (double r : {d.surface_inner, d.surface_outer}) {
      const double rr = r / 6.0;
      const double t = d.temperature_at_six / std::sqrt(rr * std::sqrt(rr));
      if (!in_lut(t))
        return set_err(err, "surface temperature at the inner or outer radius lies outside the "
                            "band table [2^-8, 2^24] K"),
               GL_ERR_DOMAIN;
    }
  }
  if (d.transport_model == GL_TRANSPORT_SLAB) {
    if (!std::isfinite(d.optical_depth) || d.optical_depth < 0.0)
      return set_err(err, "optical depth must be finite and non-negative"), GL_ERR_INVALID_ARGUMENT;
    if (!std::isfinite(d.slab_emission) || d.slab_emission < 0.0)
      return set_err(err, "slab emission must be finite and non-negative"), GL_ERR_INVALID_ARGUMENT;
    if (!std::isfinite(d.slab_temperature) || d.slab_temperature < 0.0)
      return set_err(err, "slab temperature must be finite and non-negative (0: neutral)"), GL_ERR_INVALID_ARGUMENT;
    if (blackbody && d.slab_emission > 0.0) {
      if (d.slab_temperature == 0.0)
        return set_err(err, "blackbody surface transport with nonzero emission requires a resolved "
                            "spectrum"),
               GL_ERR_DOMAIN;
      if (!in_lut(d.slab_temperature))
        return set_err(err, "slab emission temperature lies outside the band table [2^-8, 2^24] K"),
               GL_ERR_DOMAIN;
    }
    const double t = std::exp(-d.optical_depth);
    if (!f32_normal_or_zero(t) || (float)t == 0.0f || !f32_normal_or_zero(d.slab_emission))
      return set_err(err, "surface transport not representable in f32"), GL_ERR_NOT_REPRESENTABLE;
  }
  return GL_OK;
}

int prepare_scene(const gl_scene_desc& d, SceneDev* out, SceneHost* host, std::string* err) {
  std::memset(out, 0, sizeof(SceneDev));
  // Finite inputs.
  const double vals[] = {d.mass, d.spin, d.charge, d.frequency, d.vertical_fov,
                         d.escape_radius, d.horizon_offset, d.rtol, d.atol, d.initial_step,
                         d.minimum_step, d.maximum_step, d.maximum_step_per_radius,
                         d.event_tie_tolerance, d.arming_band, d.drift_limit,
                         d.subpixel[0], d.subpixel[1], d.step_safety, d.step_min_factor,
                         d.step_max_factor, d.near_critical_factor, d.branch_factor,
                         d.refine_rtol, d.refine_atol};
  for (double v : vals)
    if (!std::isfinite(v)) return set_err(err, "scene value is not finite"), GL_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < 4; ++i)
    if (!std::isfinite(d.observer_txyz[i]) || !std::isfinite(d.target_txyz[i]))
      return set_err(err, "observer or target is not finite"), GL_ERR_INVALID_ARGUMENT;
  // Reference-v1 normalization (geodesic.rs:28-41, input.rs:57-63).
  if (d.mass != 1.0) return set_err(err, "mass must be exactly 1 (M-normalized scene)"), GL_ERR_DOMAIN;
  if (std::fabs(d.frequency - 1.0) > 32.0 * 2.220446049250313e-16)
    return set_err(err, "observer frequency must be normalized to 1"), GL_ERR_DOMAIN;
  if (d.chart != 1 && d.chart != -1) return set_err(err, "chart must be +1 or -1"), GL_ERR_INVALID_ARGUMENT;
  if (d.width == 0 || d.height == 0) return set_err(err, "viewport must be nonempty"), GL_ERR_INVALID_ARGUMENT;
  if (d.width > 65536 || d.height > 65536) return set_err(err, "viewport side above 65536"), GL_ERR_CAPACITY;
  // pixel indices are 32-bit: the count must fit (the checked width x height
  // product of scientific_capture.rs:294)
  if (uint64_t(d.width) * d.height > 0xffffffffull)
    return set_err(err, "viewport above 2^32 - 1 pixels"), GL_ERR_CAPACITY;
  if (!(d.vertical_fov > 0.0 && d.vertical_fov < M_PI))
    return set_err(err, "vertical field of view must lie in (0, pi)"), GL_ERR_INVALID_ARGUMENT;
  if (!(d.subpixel[0] >= 0.0 && d.subpixel[0] <= 1.0 && d.subpixel[1] >= 0.0 && d.subpixel[1] <= 1.0))
    return set_err(err, "subpixel offsets must lie in [0, 1]"), GL_ERR_INVALID_ARGUMENT;
  if (!(d.escape_radius > 0.0)) return set_err(err, "escape radius must be positive"), GL_ERR_INVALID_ARGUMENT;
  if (!(d.rtol > 0.0 && d.atol > 0.0 && d.minimum_step > 0.0 && d.maximum_step >= d.minimum_step &&
        d.initial_step >= d.minimum_step && d.initial_step <= d.maximum_step))
    return set_err(err, "invalid step policy"), GL_ERR_INVALID_ARGUMENT;
  {
    // controller factors (policy.rs:38-66): 0 < safety <= 1, 0 < min <= 1 <= max
    const double sf = d.step_safety == 0.0 ? 0.9 : d.step_safety;
    const double mn = d.step_min_factor == 0.0 ? 0.2 : d.step_min_factor;
    const double mx = d.step_max_factor == 0.0 ? 5.0 : d.step_max_factor;
    if (!(sf > 0.0 && sf <= 1.0) || !(mn > 0.0 && mn <= 1.0) || !(mx >= 1.0 && mx <= 1.0e3))
      return set_err(err, "step controller factors must satisfy 0 < safety <= 1, 0 < min <= 1 <= max"),
             GL_ERR_INVALID_ARGUMENT;
  }
  if (d.refine_rtol < 0.0 || d.refine_atol < 0.0 ||
      (d.refine_rtol > 0.0 && !(d.refine_rtol >= 1e-15 && d.refine_rtol <= 1e-3)) ||
      (d.refine_atol > 0.0 && !(d.refine_atol >= 1e-18 && d.refine_atol <= 1e-3)))
    return set_err(err, "refine tolerances must lie in [1e-15, 1e-3] (rtol) and [1e-18, 1e-3] "
                        "(atol), or be 0 for the defaults"),
           GL_ERR_INVALID_ARGUMENT;
  if (d.near_critical_factor > 1.0e6 || d.branch_factor > 1.0e9)
    return set_err(err, "near-critical factor must not exceed 1e6, the branch factor 1e9"),
           GL_ERR_INVALID_ARGUMENT;
  {
    const double nf = d.near_critical_factor == 0.0 ? 16.0 : d.near_critical_factor;
    const double bf = d.branch_factor == 0.0 ? 4096.0 : d.branch_factor;
    if (nf > 0.0 && !(bf >= nf))
      return set_err(err, "the branch factor must be at least the near-critical factor"),
             GL_ERR_INVALID_ARGUMENT;
  }
  if (d.max_accepted_steps == 0 || d.max_consecutive_rejects == 0)
    return set_err(err, "step limits must be positive"), GL_ERR_INVALID_ARGUMENT;
  if (!(d.horizon_offset >= 0.0 && d.horizon_offset < 1.0))
    return set_err(err, "horizon offset must lie in [0, 1)"), GL_ERR_INVALID_ARGUMENT;
  if ((d.options & GL_OPT_CAPTURE_CERTIFICATE) && d.has_surface)
    return set_err(err, "the capture certificate applies to scenes without a surface: a certified "
                        "ray could still cross the disk on its way in"),
           GL_ERR_INVALID_ARGUMENT;
  if (d.has_surface) {
    if (!(d.surface_inner > 0.0 && d.surface_outer >= d.surface_inner))
      return set_err(err, "surface radii must be positive and ordered"), GL_ERR_INVALID_ARGUMENT;
    // events.rs:73-78 and input.rs:342-381: strictly inside the escape boundary.
    if (!(d.surface_outer < d.escape_radius))
      return set_err(err, "surface outer radius must lie strictly inside the escape radius"), GL_ERR_DOMAIN;
    const int st = check_surface_radiance(d, err);
    if (st != GL_OK) return st;
  }

  // Horizon and extremality, unchanged by the f32 packing (input.rs:76-90).
  int ext = extremality_f64(d.mass, d.spin, d.charge);
  float spin_f = (float)d.spin, charge_f = (float)d.charge;
  if (!std::isfinite(spin_f) || !std::isfinite(charge_f))
    return set_err(err, "spin or charge not representable in f32"), GL_ERR_NOT_REPRESENTABLE;
  if (extremality_f32(1.0f, spin_f, charge_f) != ext)
    return set_err(err, "extremality changes under the f32 packing"), GL_ERR_NOT_REPRESENTABLE;
  double r_plus = -1.0;
  if (ext == 0)
    r_plus = d.mass + std::sqrt(d.mass * d.mass - d.spin * d.spin - d.charge * d.charge);
  else if (ext == 1)
    r_plus = d.mass;

  // Stationary observer: u = d_t / sqrt(-g_tt) (observer.rs:67-104).
  Metric m;
  const double* ox = &d.observer_txyz[1];
  if (!metric_at(d, ox, &m)) return set_err(err, "metric undefined at the observer"), GL_ERR_DOMAIN;
  double g_tt = m.g[0][0];
  if (!(g_tt < 0.0)) return set_err(err, "a stationary observer requires g_tt < 0"), GL_ERR_DOMAIN;
  V4 u = {{1.0 / std::sqrt(-g_tt), 0.0, 0.0, 0.0}};
  V4 sight_seed;
  for (int i = 0; i < 4; ++i) sight_seed.v[i] = d.target_txyz[i] - d.observer_txyz[i];
  V4 sight;
  if (!orthonormalize(m, u, nullptr, 0, sight_seed, &sight))
    return set_err(err, "target does not define a direction in the observer rest frame"), GL_ERR_DOMAIN;
  V4 arrival;
  for (int i = 0; i < 4; ++i) arrival.v[i] = -sight.v[i];
  V4 up;
  int up_axis = 0;
  V4 up_seed = {{0.0, d.up_hint[0], d.up_hint[1], d.up_hint[2]}};
  if (!orthonormalize(m, u, &sight, 1, up_seed, &up)) {
    // Fallback: the coordinate axis whose normalized remainder has the
    // largest spatial component (best_coordinate_axis, observer.rs:319-352).
    double best = -1.0;
    bool found = false;
    for (int a = 1; a <= 3; ++a) {
      V4 s = {{0, 0, 0, 0}};
      s.v[a] = 1.0;
      V4 c;
      if (!orthonormalize(m, u, &sight, 1, s, &c)) continue;
      double score = std::fmax(std::fabs(c.v[1]), std::fmax(std::fabs(c.v[2]), std::fabs(c.v[3])));
      if (score >= best * (1.0 - 1.0e-9)) {  // ties go to the later axis
        best = score;
        up = c;
        up_axis = a;
        found = true;
      }
    }
    if (!found) return set_err(err, "no stable image-up axis"), GL_ERR_DOMAIN;
  }
  // Right leg: the remaining unit direction, oriented so that
  // det(u, right, up, arrival) > 0. Axes are compared by the largest
  // spatial component of their normalized remainder, the later axis winning
  // ties (Iterator::max_by). Two axes that normalize to the same direction
  // tie exactly in exact arithmetic; the one with the tiny remainder carries
  // its cancellation error into its score and, picked by that last-bit
  // accident, fails the Gram budget, so scores within a relative 1e-9 tie.
  V4 right;
  {
    V4 basis[2] = {sight, up};
    double best = -1.0;
    bool found = false;
    for (int a = 1; a <= 3; ++a) {
      V4 s = {{0, 0, 0, 0}};
      s.v[a] = 1.0;
      V4 c;
      if (!orthonormalize(m, u, basis, 2, s, &c)) continue;
      double score = std::fmax(std::fabs(c.v[1]), std::fmax(std::fabs(c.v[2]), std::fabs(c.v[3])));
      if (score >= best * (1.0 - 1.0e-9)) {  // ties go to the later axis
        best = score;
        right = c;
        found = true;
      }
    }
    if (!found) return set_err(err, "observer tetrad could not be completed"), GL_ERR_DOMAIN;
    V4 cols[4] = {u, right, up, arrival};
    if (det4(cols) < 0.0)
      for (int i = 0; i < 4; ++i) right.v[i] = -right.v[i];
  }
  // Gram check of the tetrad (observer.rs:354-376, 1e-12 budget): each
  // residual relative to the magnitude of the terms it cancels, so that a
  // frame in a strong-field metric with large components is judged by its
  // rounding, not by the size of g.
  {
    V4 b[4] = {u, right, up, arrival};
    double worst = 0.0;
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        double e = i == j ? (i == 0 ? -1.0 : 1.0) : 0.0;
        double norm = 0.0;
        const double v = dot_term_norm(m, b[i], b[j], &norm);
        worst = std::fmax(worst, std::fabs(v - e) / std::fmax(norm, 1.0));
      }
    if (!(worst <= 1e-12)) return set_err(err, "observer frame failed its Gram check"), GL_ERR_DOMAIN;
  }

  // Observer must start strictly inside the escape surface (escape armed).
  double r_obs = geometry<double>(ox[0], ox[1], ox[2], d.mass, d.spin, d.spin * d.spin,
                                  d.charge * d.charge, double(d.chart)).r;
  if (!(r_obs < d.escape_radius)) return set_err(err, "observer must lie inside the escape radius"), GL_ERR_DOMAIN;
  if (r_plus > 0.0 && !(r_obs > r_plus)) return set_err(err, "observer must lie outside the horizon"), GL_ERR_DOMAIN;

  SceneDev& s = *out;
  s.mass = 1.0f;
  s.spin = spin_f;
  s.spin2 = spin_f * spin_f;
  s.charge2 = charge_f * charge_f;
  s.chart = d.chart >= 0 ? 1.0f : -1.0f;
  s.r_horizon = (float)r_plus;
  s.r_capture = r_plus > 0.0 ? (float)(r_plus * (1.0 + d.horizon_offset)) : -1.0f;
  s.r_escape = (float)d.escape_radius;
  s.disk_in = d.has_surface ? (float)d.surface_inner : -1.0f;
  s.disk_out = d.has_surface ? (float)d.surface_outer : -1.0f;
  s.emission_model = d.has_surface ? d.emission_model : 0u;
  s.arming_band = (float)d.arming_band;
  s.rtol = (float)d.rtol;
  s.atol = (float)d.atol;
  s.h_init = (float)d.initial_step;
  s.h_min = (float)d.minimum_step;
  s.h_max = (float)d.maximum_step;
  s.h_max_per_r = (float)d.maximum_step_per_radius;
  s.safety = d.step_safety == 0.0 ? 0.9f : (float)d.step_safety;
  s.min_factor = d.step_min_factor == 0.0 ? 0.2f : (float)d.step_min_factor;
  s.max_factor = d.step_max_factor == 0.0 ? 5.0f : (float)d.step_max_factor;
  s.tie_tol = (float)d.event_tie_tolerance;
  s.drift_limit = (float)d.drift_limit;
  s.guard = (float)std::ldexp(1.0, -40);  // (r^4 + a^2 z^2) / M^4 = 2^-40 (policy.rs:3)
  s.max_accepted = d.max_accepted_steps;
  s.max_rejects = d.max_consecutive_rejects;
  s.options = d.options;
  s.reproject_interval = d.reproject_interval;
  s.obs_x = (float)ox[0];
  s.obs_y = (float)ox[1];
  s.obs_z = (float)ox[2];
  s.u_t = (float)u.v[0];
  for (int i = 0; i < 4; ++i) {
    s.right[i] = (float)right.v[i];
    s.up[i] = (float)up.v[i];
    s.arrival[i] = (float)arrival.v[i];
  }
  double thf = std::tan(0.5 * d.vertical_fov);
  s.tan_half_fov = (float)thf;
  s.aspect = (float)((double)d.width / (double)d.height);
  s.sub_x = (float)d.subpixel[0];
  s.sub_y = (float)d.subpixel[1];
  s.omega = (float)d.frequency;
  s.width = d.width;
  s.height = d.height;
  // A nonzero f64 height may round to zero in f32: keep the discrete side
  // from the f64 value (input.rs:93-109).
  s.polar_side = ox[2] > 0.0 ? 2u : (ox[2] < 0.0 ? 0u : 1u);
  s.ray_offset = 0;
  s.tiles_x = (d.width + 7) / 8;
  s.spin_d = d.spin;
  s.charge2_d = d.charge * d.charge;
  s.r_horizon_d = r_plus;
  s.r_capture_d = r_plus > 0.0 ? r_plus * (1.0 + d.horizon_offset) : -1.0;
  s.disk_in_d = d.has_surface ? d.surface_inner : -1.0;
  s.disk_out_d = d.has_surface ? d.surface_outer : -1.0;
  const bool slab = d.has_surface && d.transport_model == GL_TRANSPORT_SLAB;
  s.intensity6_d = d.has_surface ? d.intensity_at_six : 0.0;
  s.temperature6_d = d.has_surface && d.emission_model == GL_EMISSION_BLACKBODY ? d.temperature_at_six : 0.0;
  s.transmittance_d = slab ? std::exp(-d.optical_depth) : 1.0;
  s.slab_emission_d = slab ? d.slab_emission : 0.0;
  s.slab_temperature_d = slab ? d.slab_temperature : 0.0;
  s.lut = nullptr;  // set by the plan when it uploads the band table
  for (int i = 0; i < 3; ++i) s.obs_d[i] = ox[i];
  s.u_t_d = u.v[0];
  for (int i = 0; i < 4; ++i) {
    s.right_d[i] = right.v[i];
    s.up_d[i] = up.v[i];
    s.arrival_d[i] = arrival.v[i];
  }
  s.tan_half_fov_d = thf;
  s.aspect_d = (double)d.width / (double)d.height;
  s.sub_x_d = d.subpixel[0];
  s.sub_y_d = d.subpixel[1];
  s.r_observer_d = r_obs;
  s.phi_observer_d = std::atan2(ox[1], ox[0]);
  s.near_critical_factor = d.near_critical_factor == 0.0 ? 16.0f
                           : (d.near_critical_factor < 0.0 ? -1.0f : (float)d.near_critical_factor);
  s.branch_factor = s.near_critical_factor < 0.0f ? -1.0f
                     : (d.branch_factor == 0.0 ? 4096.0f : (float)d.branch_factor);
  s.refine_rtol = d.refine_rtol == 0.0 ? 1e-12f : (float)d.refine_rtol;
  s.refine_atol = d.refine_atol == 0.0 ? 1e-14f : (float)d.refine_atol;
  if (!std::isfinite(s.tan_half_fov) || !std::isfinite(s.u_t))
    return set_err(err, "scene not representable in f32"), GL_ERR_NOT_REPRESENTABLE;

  if (host) {
    host->r_plus = r_plus;
    host->extremality = ext;
    host->g_tt = g_tt;
    for (int i = 0; i < 4; ++i) {
      host->u[i] = u.v[i];
      host->right[i] = right.v[i];
      host->up[i] = up.v[i];
      host->arrival[i] = arrival.v[i];
    }
    host->tan_half_fov = thf;
    host->up_axis = up_axis;
    host->desc = d;
  }
  return GL_OK;
}

}  // namespace gl
