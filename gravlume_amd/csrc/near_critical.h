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
  const T p = (T(2) * E * c - B) / (T(2) * E2);
  const T s = B / (T(2) * E2);
  T r = r_plus;
  if (p < T(0) && T(4) * p * p * p + T(27) * s * s < T(0)) {
    // largest root of r^3 + p r + s: Newton from 2 sqrt(-p/3), above it, where
    // the cubic is positive, increasing and convex, so the iterates decrease
    // monotonically onto the root
    T x = T(2) * gl_sqrt<T>(-p / T(3));
    for (int it = 0; it < 60; ++it) {
      const T f = (x * x + p) * x + s;
      const T df = T(3) * x * x + p;
      if (!(df > T(0))) break;
      const T nx = x - f / df;
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
