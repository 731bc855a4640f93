// Oracle (test infrastructure only): stationary observer tetrad, perspective
// view and initial photon momentum, f64. Restates
// crates/gravlume-domain/src/{observer,scene,view}.rs.
#include <cmath>

#include "oracle.hpp"

namespace orc {
namespace {

constexpr double kBasisTolerance = 1.0e-12;       // observer.rs:7
constexpr double kInitialNullTolerance = 2.0e-12;  // scene.rs:8

struct FV {
  double v[4];
};

FV fv_add(const FV& a, const FV& b) {
  FV r;
  for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] + b.v[i];
  return r;
}
FV fv_sub(const FV& a, const FV& b) {
  FV r;
  for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] - b.v[i];
  return r;
}
FV fv_scale(const FV& a, double s) {
  FV r;
  for (int i = 0; i < 4; ++i) r.v[i] = a.v[i] * s;
  return r;
}
bool fv_finite(const FV& a) {
  for (double x : a.v)
    if (!std::isfinite(x)) return false;
  return true;
}

// metric_dot_and_term_norm (spacetime.rs:463-483): row-major sum of g_rc l_r r_c.
bool metric_dot(const double g[4][4], const FV& l, const FV& r, double* dot, double* norm) {
  double s = 0.0, n = 0.0;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      double t = g[a][b] * l.v[a] * r.v[b];
      s += t;
      n += std::fabs(t);
    }
  if (!std::isfinite(s) || !std::isfinite(n)) return false;
  *dot = s;
  if (norm) *norm = n;
  return true;
}

// project_and_normalize (observer.rs:289-317).
bool project_and_normalize(const double g[4][4], const FV& u, const FV& seed, const FV* ortho,
                           int n_ortho, FV* out) {
  if (!fv_finite(seed)) return false;
  double uds;
  if (!metric_dot(g, u, seed, &uds, nullptr)) return false;
  FV p = fv_add(seed, fv_scale(u, uds));
  for (int i = 0; i < n_ortho; ++i) {
    double c;
    if (!metric_dot(g, p, ortho[i], &c, nullptr)) return false;
    p = fv_sub(p, fv_scale(ortho[i], c));
  }
  double n2;
  if (!metric_dot(g, p, p, &n2, nullptr)) return false;
  if (!std::isfinite(n2) || n2 <= kBasisTolerance * kBasisTolerance) return false;
  FV r = fv_scale(p, 1.0 / std::sqrt(n2));
  if (!fv_finite(r)) return false;
  *out = r;
  return true;
}

// best_coordinate_axis (observer.rs:319-352): the candidate with the largest
// spatial max-norm; on ties the later axis wins (Iterator::max_by). Two axes
// whose remainders normalize to the same direction tie exactly in exact
// arithmetic, but the one with the tiny remainder (an axis nearly inside the
// span already projected out) carries its cancellation error into the score,
// which then decides the pick by a last-bit accident and, when it goes to
// that axis, the frame fails the 1e-12 Gram budget. Scores within a relative
// 1e-9 are therefore taken as the tie they are in exact arithmetic.
bool best_coordinate_axis(const double g[4][4], const FV& u, const FV* ortho, int n_ortho,
                          FV* out) {
  bool found = false;
  double best = -1.0;
  for (int axis = 1; axis <= 3; ++axis) {
    FV seed = {{0.0, 0.0, 0.0, 0.0}};
    seed.v[axis] = 1.0;
    FV cand;
    if (!project_and_normalize(g, u, seed, ortho, n_ortho, &cand)) continue;
    double score = 0.0;
    for (int i = 1; i < 4; ++i) score = std::fmax(score, std::fabs(cand.v[i]));
    if (!found || score >= best * (1.0 - 1.0e-9)) {
      best = score;
      *out = cand;
      found = true;
    }
  }
  return found;
}

// determinant4 (observer.rs:378-408): partial-pivot elimination on the matrix
// whose columns are the four basis vectors.
double determinant4(const FV cols[4]) {
  double m[4][4];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) m[r][c] = cols[c].v[r];
  double det = 1.0;
  for (int pc = 0; pc < 4; ++pc) {
    int pr = pc;
    for (int r = pc + 1; r < 4; ++r)
      if (std::fabs(m[r][pc]) >= std::fabs(m[pr][pc])) pr = r;
    if (m[pr][pc] == 0.0) return 0.0;
    if (pr != pc) {
      for (int c = 0; c < 4; ++c) std::swap(m[pr][c], m[pc][c]);
      det = -det;
    }
    double piv = m[pc][pc];
    det *= piv;
    for (int r = pc + 1; r < 4; ++r) {
      double f = m[r][pc] / piv;
      for (int c = pc + 1; c < 4; ++c) m[r][c] = std::fma(f, -m[pc][c], m[r][c]);
    }
  }
  return det;
}

// gram_residual (observer.rs:354-376).
double gram_residual(const double g[4][4], const FV basis[4]) {
  double res = 0.0;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double dot, norm;
      if (!metric_dot(g, basis[r], basis[c], &dot, &norm)) return INFINITY;
      double expected = r == c ? (r == 0 ? -1.0 : 1.0) : 0.0;
      res = std::fmax(res, std::fabs(dot - expected) / std::fmax(norm, 1.0));
    }
  return res;
}

}  // namespace

// PhysicalScene::new + StationaryObserverInput::build + construct_frame +
// PerspectiveView::new (scene.rs:112-143, observer.rs:67-183, view.rs:42-69).
// Returns 0 on success, a nonzero validation code otherwise.
int build_observation(double mass, double spin, double charge, double chart,
                      const double event[4], const double target[4], const double up_hint[3],
                      double omega, uint32_t width, uint32_t height, double fov,
                      Observation* out) {
  Observation o;
  if (make_spacetime(mass, spin, charge, chart, &o.st)) return 1;
  for (int i = 0; i < 4; ++i)
    if (!std::isfinite(event[i]) || !std::isfinite(target[i])) return 2;
  for (int i = 0; i < 3; ++i)
    if (!std::isfinite(up_hint[i])) return 2;
  if (!std::isfinite(omega) || omega <= 0.0) return 3;
  if (width == 0 || height == 0) return 4;
  if (!(fov > 0.0 && fov < M_PI)) return 4;
  double thf = std::tan(fov * 0.5);
  if (!std::isfinite(thf)) return 4;
  Geometry g;
  if (geometry(o.st, &event[1], &g)) return 5;
  double g_tt = -1.0 + g.f;
  if (g_tt >= 0.0) return 6;  // NonStationaryObserver
  double metric[4][4];
  metric_covariant(g, metric);
  FV u = {{1.0 / std::sqrt(-g_tt), 0.0, 0.0, 0.0}};
  FV target_seed;
  for (int i = 0; i < 4; ++i) target_seed.v[i] = target[i] - event[i];
  FV sight;
  if (!project_and_normalize(metric, u, target_seed, nullptr, 0, &sight)) return 7;
  FV arrival = fv_scale(sight, -1.0);
  FV up_seed = {{0.0, up_hint[0], up_hint[1], up_hint[2]}};
  FV up;
  int used_fallback = 0;
  if (!project_and_normalize(metric, u, up_seed, &sight, 1, &up)) {
    if (!best_coordinate_axis(metric, u, &sight, 1, &up)) return 8;
    used_fallback = 1;
  }
  FV ortho2[2] = {sight, up};
  FV right;
  if (!best_coordinate_axis(metric, u, ortho2, 2, &right)) return 9;
  FV cols[4] = {u, right, up, arrival};
  double orient = determinant4(cols);
  if (orient < 0.0) {
    right = fv_scale(right, -1.0);
    orient = -orient;
  }
  FV basis[4] = {u, right, up, arrival};
  double gram = gram_residual(metric, basis);
  if (!std::isfinite(gram) || gram > kBasisTolerance || !std::isfinite(orient) || orient <= 0.0)
    return 10;
  for (int i = 0; i < 4; ++i) {
    o.event[i] = event[i];
    o.u[i] = u.v[i];
    o.right[i] = right.v[i];
    o.up[i] = up.v[i];
    o.arrival[i] = arrival.v[i];
    for (int j = 0; j < 4; ++j) o.metric[i][j] = metric[i][j];
  }
  o.gram_residual = gram;
  o.orientation = orient;
  o.used_up_fallback = used_fallback;
  o.g_tt = g_tt;
  o.omega = omega;
  o.width = width;
  o.height = height;
  o.fov = fov;
  o.tan_half_fov = thf;
  *out = o;
  return 0;
}

// PerspectiveView::sight_plane (view.rs:107-123).
int sight_plane(const Observation& obs, uint32_t px, uint32_t py, double subx, double suby,
                double* sx, double* sy) {
  if (px >= obs.width || py >= obs.height) return 1;
  if (!(subx >= 0.0 && subx <= 1.0) || !(suby >= 0.0 && suby <= 1.0)) return 1;
  double w = static_cast<double>(obs.width), h = static_cast<double>(obs.height);
  double nx = 2.0 * (static_cast<double>(px) + subx) / w - 1.0;
  double ny = 1.0 - 2.0 * (static_cast<double>(py) + suby) / h;
  double aspect = w / h;
  *sx = aspect * obs.tan_half_fov * nx;
  *sy = obs.tan_half_fov * ny;
  return 0;
}

// Observation::initial_ray (scene.rs:211-271) for a given sight-plane point.
int initial_ray_sight(const Observation& obs, double sx, double sy, S8* state, double sight[4],
                      double* frequency, double* null_residual) {
  double norm = 1.0 / std::hypot(1.0, std::hypot(sx, sy));
  FV right, up, arrival, u;
  for (int i = 0; i < 4; ++i) {
    right.v[i] = obs.right[i];
    up.v[i] = obs.up[i];
    arrival.v[i] = obs.arrival[i];
    u.v[i] = obs.u[i];
  }
  FV d = fv_scale(fv_sub(fv_add(fv_scale(right, sx), fv_scale(up, sy)), arrival), norm);
  FV n = fv_scale(d, -1.0);
  FV pc = fv_scale(fv_add(u, n), obs.omega);
  if (!fv_finite(pc)) return 1;
  double pl[4];
  for (int r = 0; r < 4; ++r) {
    double s = 0.0;
    for (int c = 0; c < 4; ++c) s += obs.metric[r][c] * pc.v[c];
    pl[r] = s;
  }
  double nr = normalized_quadratic_form_residual(obs.metric, pc.v);
  double freq = 0.0;
  for (int i = 0; i < 4; ++i) freq += pl[i] * u.v[i];
  freq = -freq;
  for (double v : pl)
    if (!std::isfinite(v)) return 1;
  if (!std::isfinite(freq) || !std::isfinite(nr)) return 1;
  if (freq <= 0.0 || nr > kInitialNullTolerance) return 2;
  for (int i = 0; i < 4; ++i) {
    (*state)[i] = obs.event[i];
    (*state)[4 + i] = pl[i];
    if (sight) sight[i] = d.v[i];
  }
  if (frequency) *frequency = freq;
  if (null_residual) *null_residual = nr;
  return 0;
}

}  // namespace orc
