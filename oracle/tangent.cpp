// Forward-mode tangents of the surface frequency ratio g on the f64 oracle.
//
// TEST INFRASTRUCTURE ONLY (the checker of the HIP adjoint in
// gravlume_amd/csrc/gl_vjp.hip); the product never calls it.
//
// For one sight-plane point (s_x, s_y) of an observation at parameters
// theta = (a, Q, inclination), with the observer at fixed oblate radius and
// azimuth, this integrates the trace together with its variational equations
//     y' = F(y; a, Q),   Y_k' = (dF/dy) Y_k + dF/dtheta_k,   k = 0..4,
// over (a, Q, inclination, s_x, s_y), with an adaptive f64 Dormand-Prince
// 5(4) (Dormand & Prince 1980) whose error control covers y and the Y_k,
// until the backward ray crosses the equator inside the disk annulus. The
// Jacobian-vector products (dF/dy) Y_k are symmetric differences of
// hamiltonian_rhs (spacetime.cpp) along Y_k at a relative displacement of
// 1e-6, and dF/da, dF/dQ symmetric differences at 1e-6; the initial tangents
// Y_k(0) are symmetric differences of the observation construction
// (build_observation + initial_ray_sight). At the crossing lambda*, where
// z(lambda*) = 0,
//     dy_T/dtheta_k = Y_k(lambda*) - F(y_T) Y_k,z(lambda*) / F_z(y_T),
// and dg/dtheta_k follows by the chain rule through observable_at
// (surface.rs:75-146 restated in observer.cpp), whose partial derivatives
// are symmetric differences at 1e-7.
#include <cmath>
#include <cstring>

#include "oracle.hpp"

namespace orc {

namespace {

constexpr int kP = 5;              // tangent directions
constexpr int kZ = 8 * (1 + kP);   // augmented state size

const double TA[7][7] = {
    {0, 0, 0, 0, 0, 0, 0},
    {1.0 / 5.0, 0, 0, 0, 0, 0, 0},
    {3.0 / 40.0, 9.0 / 40.0, 0, 0, 0, 0, 0},
    {44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0, 0, 0, 0, 0},
    {19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0, 0, 0, 0},
    {9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0, 0, 0},
    {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0, 0},
};
const double TB5[7] = {35.0 / 384.0, 0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0,
                       11.0 / 84.0, 0};
const double TB4[7] = {5179.0 / 57600.0, 0, 7571.0 / 16695.0, 393.0 / 640.0,
                       -92097.0 / 339200.0, 187.0 / 2100.0, 1.0 / 40.0};

struct Setup {
  double a, q, incl, chart, r_obs;
  uint32_t width, height;
  double fov;
};

void observer_event(const Setup& s, double a, double incl, double ev[4]) {
  double st = std::sin(incl), ct = std::cos(incl);
  double cs = (s.chart >= 0.0 ? 1.0 : -1.0) * a;
  ev[0] = 0.0;
  ev[1] = s.r_obs * st;  // azimuth 0: (r cos 0 - cs sin 0) sin(incl)
  ev[2] = cs * st;
  ev[3] = s.r_obs * ct;
}

int initial_state(const Setup& s, double a, double q, double incl, double sx, double sy, S8* y) {
  double ev[4], target[4] = {0, 0, 0, 0}, up[3] = {0, 0, 1};
  observer_event(s, a, incl, ev);
  Observation obs;
  int e = build_observation(1.0, a, q, s.chart, ev, target, up, 1.0, s.width, s.height, s.fov,
                            &obs);
  if (e) return e;
  double sight[4], fn, nres;
  return initial_ray_sight(obs, sx, sy, y, sight, &fn, &nres);
}

int rhs_at(double a, double q, double chart, const S8& y, S8* f) {
  Spacetime st;
  int e = make_spacetime(1.0, a, q, chart, &st);
  if (e) return e;
  return hamiltonian_rhs(st, y, f);
}

// Augmented right-hand side.
int aug_rhs(const Setup& s, const double z[kZ], double out[kZ]) {
  S8 y, f;
  std::memcpy(y.data(), z, sizeof(double) * 8);
  if (rhs_at(s.a, s.q, s.chart, y, &f)) return 1;
  std::memcpy(out, f.data(), sizeof(double) * 8);
  double ny = 0.0;
  for (int c = 0; c < 8; ++c) ny = std::fmax(ny, std::fabs(y[c]));
  for (int k = 0; k < kP; ++k) {
    const double* Y = z + 8 * (1 + k);
    double nY = 0.0;
    for (int c = 0; c < 8; ++c) nY = std::fmax(nY, std::fabs(Y[c]));
    double* o = out + 8 * (1 + k);
    for (int c = 0; c < 8; ++c) o[c] = 0.0;
    if (nY > 0.0) {
      const double eps = 1e-6 * std::fmax(1.0, ny) / nY;
      S8 yp, ym, fp, fm;
      for (int c = 0; c < 8; ++c) {
        yp[c] = y[c] + eps * Y[c];
        ym[c] = y[c] - eps * Y[c];
      }
      if (rhs_at(s.a, s.q, s.chart, yp, &fp) || rhs_at(s.a, s.q, s.chart, ym, &fm)) return 1;
      for (int c = 0; c < 8; ++c) o[c] = (fp[c] - fm[c]) / (2.0 * eps);
    }
    if (k < 2) {  // explicit dependence on a and Q
      const double d = 1e-6;
      S8 fp, fm;
      double ap = s.a, am = s.a, qp = s.q, qm = s.q;
      if (k == 0) ap += d, am -= d;
      else qp += d, qm -= d;
      if (rhs_at(ap, qp, s.chart, y, &fp) || rhs_at(am, qm, s.chart, y, &fm)) return 1;
      for (int c = 0; c < 8; ++c) o[c] += (fp[c] - fm[c]) / (2.0 * d);
    }
  }
  return 0;
}

// One DP5 step of the augmented system; err = scaled error norm.
int aug_step(const Setup& s, const double z[kZ], double h, double rtol, double atol,
             double out[kZ], double* err) {
  double K[7][kZ];
  if (aug_rhs(s, z, K[0])) return 1;
  double tmp[kZ];
  for (int st = 1; st < 7; ++st) {
    for (int c = 0; c < kZ; ++c) {
      double acc = 0.0;
      for (int j = 0; j < st; ++j) acc += TA[st][j] * K[j][c];
      tmp[c] = z[c] + h * acc;
    }
    if (aug_rhs(s, tmp, K[st])) return 1;
  }
  double e = 0.0;
  for (int c = 0; c < kZ; ++c) {
    double y5 = 0.0, y4 = 0.0;
    for (int j = 0; j < 7; ++j) {
      y5 += TB5[j] * K[j][c];
      y4 += TB4[j] * K[j][c];
    }
    out[c] = z[c] + h * y5;
    const double sc = atol + rtol * std::fmax(std::fabs(z[c]), std::fabs(out[c]));
    e = std::fmax(e, std::fabs(h * (y5 - y4)) / sc);
  }
  *err = e;
  return 0;
}

double radius_of(const Setup& s, const double* xyz) {
  Spacetime st;
  make_spacetime(1.0, s.a, s.q, s.chart, &st);
  Geometry g;
  if (geometry(st, xyz, &g)) return -1.0;
  return g.radius;
}

double g_of(const Setup& s, double a, double q, const S8& y, double inner, double outer) {
  Spacetime st;
  if (make_spacetime(1.0, a, q, s.chart, &st)) return NAN;
  Emitter em;
  Outcome o;
  std::memset(&o, 0, sizeof(o));
  if (observable_at(st, em, inner, outer, y, 1.0, &o)) return NAN;
  return o.frequency_ratio;
}

}  // namespace

int tangent_frequency_ratio(double a, double q, double incl, double chart, double r_obs,
                            uint32_t width, uint32_t height, double fov, double sx, double sy,
                            double inner, double outer, double escape, double rtol, double* g,
                            double dg[5]) {
  Setup s{a, q, incl, chart, r_obs, width, height, fov};
  double z[kZ];
  S8 y0;
  if (initial_state(s, a, q, incl, sx, sy, &y0)) return 1;
  std::memcpy(z, y0.data(), sizeof(double) * 8);
  for (int k = 0; k < kP; ++k) {
    const double d = 1e-6;
    double p[5] = {a, q, incl, sx, sy}, m[5] = {a, q, incl, sx, sy};
    p[k] += d;
    m[k] -= d;
    S8 yp, ym;
    if (initial_state(s, p[0], p[1], p[2], p[3], p[4], &yp) ||
        initial_state(s, m[0], m[1], m[2], m[3], m[4], &ym))
      return 2;
    for (int c = 0; c < 8; ++c) z[8 * (1 + k) + c] = (yp[c] - ym[c]) / (2.0 * d);
  }
  Spacetime st;
  make_spacetime(1.0, a, q, chart, &st);
  const double r_capture = st.outer_horizon * (1.0 + 1e-6);
  double h = -0.05 * r_obs;  // backward ray
  const double atol = rtol * 1e-2;
  for (int it = 0; it < 200000; ++it) {
    double z1[kZ], e;
    if (aug_step(s, z, h, rtol, atol, z1, &e)) return 3;
    if (!(e <= 1.0)) {
      h *= std::fmax(0.2, 0.9 * std::pow(e, -0.2));
      continue;
    }
    const double r1 = radius_of(s, z1 + 1);
    if (r1 < 0.0 || r1 <= r_capture || r1 >= escape) return 4;  // not a surface hit
    const bool cross = (z[3] > 0.0) != (z1[3] > 0.0);
    if (cross) {
      // equator crossing inside the step: Newton / secant on the fraction
      double lo = 0.0, hi = 1.0, zlo = z[3], th = zlo / (zlo - z1[3]);
      double zt[kZ];
      for (int k2 = 0; k2 < 60; ++k2) {
        if (aug_step(s, z, th * h, rtol, atol, zt, &e)) return 3;
        const double v = zt[3];
        if (std::fabs(v) <= 1e-15 * std::fmax(1.0, std::fabs(z[3]))) break;
        if ((v > 0.0) == (zlo > 0.0)) lo = th;
        else hi = th;
        S8 yt, ft;
        std::memcpy(yt.data(), zt, sizeof(double) * 8);
        if (rhs_at(a, q, chart, yt, &ft)) return 3;
        double nt = th - v / (ft[3] * h);
        if (!(nt > lo && nt < hi)) nt = 0.5 * (lo + hi);
        th = nt;
        if (hi - lo < 1e-16) break;
      }
      const double rT = radius_of(s, zt + 1);
      if (rT >= inner && rT <= outer) {
        S8 yT, fT;
        std::memcpy(yT.data(), zt, sizeof(double) * 8);
        if (rhs_at(a, q, chart, yT, &fT)) return 3;
        const double g0 = g_of(s, a, q, yT, inner, outer);
        if (!std::isfinite(g0)) return 5;
        // partial derivatives of g(y_T; a, Q)
        double dgdy[8], dga = 0.0, dgq = 0.0;
        for (int c = 0; c < 8; ++c) {
          const double d = 1e-7 * std::fmax(1.0, std::fabs(yT[c]));
          S8 yp = yT, ym = yT;
          yp[c] += d;
          ym[c] -= d;
          dgdy[c] = (g_of(s, a, q, yp, inner, outer) - g_of(s, a, q, ym, inner, outer)) / (2.0 * d);
        }
        dga = (g_of(s, a + 1e-7, q, yT, inner, outer) - g_of(s, a - 1e-7, q, yT, inner, outer)) / 2e-7;
        dgq = (g_of(s, a, q + 1e-7, yT, inner, outer) - g_of(s, a, q - 1e-7, yT, inner, outer)) / 2e-7;
        *g = g0;
        for (int k = 0; k < kP; ++k) {
          const double* Y = zt + 8 * (1 + k);
          const double shift = Y[3] / fT[3];
          double v = 0.0;
          for (int c = 0; c < 8; ++c) v += dgdy[c] * (Y[c] - fT[c] * shift);
          if (k == 0) v += dga;
          if (k == 1) v += dgq;
          dg[k] = v;
        }
        return std::isfinite(dg[0]) ? 0 : 5;
      }
    }
    std::memcpy(z, z1, sizeof(z));
    const double f = std::fmin(5.0, std::fmax(0.2, 0.9 * std::pow(std::fmax(e, 1e-10), -0.2)));
    h *= f;
    const double rr = radius_of(s, z + 1);
    const double hmax = 0.25 * rr;
    if (std::fabs(h) > hmax) h = -hmax;
  }
  return 6;
}

}  // namespace orc
