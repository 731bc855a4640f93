// Near-critical test of a Kerr-Newman null geodesic (M = 1).
//
// With the constants of motion E = -p_t, L = L_z and Carter's Q, the radial
// motion obeys (Sigma dr/dmino)^2 = R(r) with the radial potential
//     R(r) = (E (r^2 + a^2) - a L)^2 - Delta(r) ((L - a E)^2 + Q),
//     Delta(r) = r^2 - 2 r + a^2 + q^2,
// (the reference's docs/physics.md:320-335 and the capture certificate's
// polynomial, docs/research/gpu-geodesic-acceleration.md:182-189). A ray
// coming in from r_start turns at the largest root of R below r_start and is
// captured when R > 0 on [r_+, r_start]. The two cases meet on the critical
// curve, where R has a double root at the minimum r_m of R over
// [r_+, r_start]; the ray's distance from it is the relative margin
//     mu = R(r_m) / ((E (r_m^2 + a^2) - a L)^2 + Delta(r_m) B),  B = (L - aE)^2 + Q,
// the potential against the magnitude of the two terms it is the
// difference of. A perturbation of the constants by a relative eps moves
// R(r_m) by about eps times that magnitude, so a trace whose constants have
// drifted by eps (or started eps away from the exact ray) settles the
// capture/scatter decision only when |mu| clearly exceeds eps; below, the
// class, the winding and the time spent near r_m follow the trace's error
// rather than the ray (docs/validation.md:287: such a sample stays
// Uncertain unless a refined trace settles it).
//
// R(r) = E^2 r^4 + (2 E c - B) r^2 + 2 B r + c^2 - (a^2 + q^2) B with
// c = a (a E - L) has no cubic term, so R' = 0 is the depressed cubic
// r^3 + p r + s = 0 (p = (2 E c - B) / (2 E^2), s = B / (2 E^2) >= 0). With
// three real roots the largest is R's rightmost local minimum; with one, that
// root is negative and R increases on r > 0, so the minimum over [r_+, ...] is
// at r_+.
#pragma once

#include "ks_math.h"

namespace gl {

template <class T>
struct RadialMargin {
  T mu;   // relative margin (> 0 capture side, < 0 scatter side)
  T r_m;  // radius of the minimum of R over [r_+, r_start]
  bool ok;
};

template <class T>
GL_HD RadialMargin<T> radial_margin(T E, T L, T Q, T spin, T charge2, T r_plus, T r_start) {
  RadialMargin<T> m;
  m.mu = T(0);
  m.r_m = r_plus;
  m.ok = false;
  if (!(E > T(0)) || !(r_plus > T(0)) || !(r_start > r_plus)) return m;
  const T a = spin, a2 = spin * spin;
  const T c = a * (a * E - L);
  const T lm = L - a * E;
  const T B = lm * lm + Q;
  const T E2 = E * E;
  const T inv2e2 = T(1) / (T(2) * E2);
  const T p = (T(2) * E * c - B) * inv2e2;
  const T s = B * inv2e2;
  T r = r_plus;
  if (p < T(0) && T(4) * p * p * p + T(27) * s * s < T(0)) {
    // largest root of r^3 + p r + s: Newton from 2 sqrt(-p/3), above it, where
    // the cubic is positive, increasing and convex, so the iterates decrease
    // monotonically onto the root
    T x = T(2) * gl_sqrt<T>(p * T(-1.0 / 3.0));
    for (int it = 0; it < 60; ++it) {
      const T f = (x * x + p) * x + s;
      const T df = T(3) * x * x + p;
      if (!(df > T(0))) break;
      const T nx = x - f * gl_rcp<T>(df);
      if (!(nx < x)) break;
      x = nx;
    }
    r = x;
  }
  if (r < r_plus) r = r_plus;
  if (r > r_start) r = r_start;
  const T A = E * (r * r + a2) - a * L;
  const T D = r * r - T(2) * r + a2 + charge2;
  const T big_r = A * A - D * B;
  const T scale = A * A + gl_abs<T>(D * B);
  if (!(scale > T(0))) return m;
  m.mu = big_r / scale;
  m.r_m = r;
  m.ok = m.mu == m.mu;
  return m;
}

// Relative change of the constants (L_z, Carter) and the null residual when
// the f64 state y (t, x, y, z, p_t, p_x, p_y, p_z) is rounded to the f32
// state the trace starts from (the position keeps its rounding as a Kahan
// term, so only the momentum and E are rounded).
GL_HD double f32_start_error(const double* y, double spin, double charge2, double chart) {
  const double x = y[1], yy = y[2], z = y[3];
  const double E = -y[4], px = y[5], py = y[6], pz = y[7];
  const double Ef = double(float(E)), pxf = double(float(px)), pyf = double(float(py)),
               pzf = double(float(pz));
  const Geom<double> g = geometry<double>(x, yy, z, 1.0, spin, spin * spin, charge2, chart);
  if (!g.ok) return 1.0;
  const double L = x * py - yy * px, Lf = x * pyf - yy * pxf;
  const double Q = carter<double>(g, x, yy, z, px, py, pz, E, spin * spin);
  const double Qf = carter<double>(g, x, yy, z, pxf, pyf, pzf, Ef, spin * spin);
  const Deriv<double> d = rhs_from_geometry<double>(g, x, yy, z, pxf, pyf, pzf, Ef, spin, chart);
  const double nr = null_residual<double>(d, pxf, pyf, pzf, Ef);
  const double el = gl_abs<double>(Lf - L) / (gl_abs<double>(L) > 1.0 ? gl_abs<double>(L) : 1.0);
  const double eq = gl_abs<double>(Qf - Q) / (gl_abs<double>(Q) > 1.0 ? gl_abs<double>(Q) : 1.0);
  double e = el > eq ? el : eq;
  return e > nr ? e : nr;
}

// The margin of the exact ray from f64 initial state y, with the scene's
// spacetime (no horizon: not near-critical).
// An upper bound of f32_start_error(y) from first-order error propagation:
// rounding E and p to binary32 moves each by at most u = 2^-24 of itself, so
//   |dL| <= u (|x p_y| + |y p_x|),
//   |dQ| <= 2 u (cz^2 D (p_x^2 + p_y^2) + cz^2 a^2 E^2 + 2 |z| (|x p_x| + |y p_y|) |p_z|
//               + r^2 (x^2 + y^2) p_z^2 / D)   (Carter's terms, quadratic in p),
// and the null contraction -E t' + p . x' of the rounded momentum is 2 H
// there (Euler: p . dH/dp = 2 H for the quadratic H), which exceeds the
// state's own 2 H by at most 2 u times the contraction's term norm, so the
// null residual is at most |2 H| + 2 u, with 2 H = -E^2 + |p|^2 - f q^2 and
// q = E + l . p (zero for a null state). A 1% margin covers the
// second-order terms and the f64 evaluation; the absolute 1e-30 covers
// binary32 subnormals. When |mu| exceeds the branch
// factor times this bound, the start error cannot flag the ray and need not
// be evaluated.
GL_HD double start_error_bound_from(const Geom<double>& g, double x, double yy, double z,
                                    double E, double px, double py, double pz, double L,
                                    double Q, double a2) {
  const double u = 5.9604644775390625e-08;  // 2^-24
  const double D = g.r * g.r + a2;
  const double cz = z * g.rinv;
  const double axpx = gl_abs<double>(x * px), aypy = gl_abs<double>(yy * py);
  const double tq = cz * cz * D * (px * px + py * py) + cz * cz * a2 * E * E +
                    2.0 * gl_abs<double>(z) * (axpx + aypy) * gl_abs<double>(pz) +
                    g.r * g.r * (x * x + yy * yy) * g.kinv * pz * pz;
  const double el = u * (gl_abs<double>(x * py) + gl_abs<double>(yy * px)) /
                    (gl_abs<double>(L) > 1.0 ? gl_abs<double>(L) : 1.0);
  const double eq = 2.0 * u * tq / (gl_abs<double>(Q) > 1.0 ? gl_abs<double>(Q) : 1.0);
  const double q = E + g.lx * px + g.ly * py + g.lz * pz;
  const double h2 = -E * E + px * px + py * py + pz * pz - g.f * q * q;
  const double scale2 = E * E + px * px + py * py + pz * pz + gl_abs<double>(g.f) * q * q;
  const double en = gl_abs<double>(h2) + 2.0 * u + 1.0e-14 * scale2;
  double e = el > eq ? el : eq;
  if (e < en) e = en;
  return 1.01 * e + 1.0e-30;
}

GL_HD double f32_start_error_bound(const double* y, double spin, double charge2, double chart) {
  const double x = y[1], yy = y[2], z = y[3];
  const double E = -y[4], px = y[5], py = y[6], pz = y[7];
  const double a2 = spin * spin;
  const Geom<double> g = geometry<double>(x, yy, z, 1.0, spin, a2, charge2, chart);
  if (!g.ok) return 1.0;
  const double L = x * py - yy * px;
  const double Q = carter<double>(g, x, yy, z, px, py, pz, E, a2);
  return start_error_bound_from(g, x, yy, z, E, px, py, pz, L, Q, a2);
}

// The margin of state y with the bound above, from one geometry evaluation.
struct MarginAndBound {
  RadialMargin<double> m;
  double bound;
};

GL_HD MarginAndBound state_margin_and_bound(const double* y, double spin, double charge2,
                                            double chart, double r_plus) {
  const double x = y[1], yy = y[2], z = y[3];
  const double E = -y[4], px = y[5], py = y[6], pz = y[7];
  const double a2 = spin * spin;
  const Geom<double> g = geometry<double>(x, yy, z, 1.0, spin, a2, charge2, chart);
  MarginAndBound r;
  r.m.mu = 0.0;
  r.m.r_m = r_plus;
  r.m.ok = false;
  r.bound = 1.0;
  if (!g.ok) return r;
  const double L = x * py - yy * px;
  const double Q = carter<double>(g, x, yy, z, px, py, pz, E, a2);
  r.m = radial_margin<double>(E, L, Q, spin, charge2, r_plus, g.r);
  r.bound = start_error_bound_from(g, x, yy, z, E, px, py, pz, L, Q, a2);
  return r;
}

GL_HD RadialMargin<double> state_margin(const double* y, double spin, double charge2, double chart,
                                        double r_plus) {
  const double x = y[1], yy = y[2], z = y[3];
  const double E = -y[4], px = y[5], py = y[6], pz = y[7];
  const Geom<double> g = geometry<double>(x, yy, z, 1.0, spin, spin * spin, charge2, chart);
  RadialMargin<double> m;
  m.mu = 0.0;
  m.r_m = r_plus;
  m.ok = false;
  if (!g.ok) return m;
  const double L = x * py - yy * px;
  const double Q = carter<double>(g, x, yy, z, px, py, pz, E, spin * spin);
  return radial_margin<double>(E, L, Q, spin, charge2, r_plus, g.r);
}

}  // namespace gl
