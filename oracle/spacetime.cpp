// Oracle (test infrastructure only): Kerr-Newman spacetime in Cartesian
// Kerr-Schild coordinates, f64. Restates crates/gravlume-domain/src/spacetime.rs,
// spacetime/exact_binary.rs and numerics.rs.
#include <cmath>
#include <limits>
#include <vector>

#include "oracle.hpp"

namespace orc {
namespace {

bool finite(double v) { return std::isfinite(v); }

// Exact sign of M^2 - (a^2 + Q^2) and its rounded value
// (restates ExactBinary, spacetime/exact_binary.rs:17-107, by a different
// method). Each square is split exactly into hi + lo with one fma
// (Dekker's two-product); the six terms are summed into a nonoverlapping
// floating-point expansion with two-sum (Shewchuk 1997, "Adaptive Precision
// Floating-Point Arithmetic", grow-expansion). The sign of an expansion is
// the sign of its largest nonzero component, which makes the extremality
// class exact. The value is rounded to within one ulp (the reference rounds
// half-even; both agree to the fixture tolerance of 2e-13 on r_+).
// Inputs are first scaled by a power of two so the largest magnitude lies in
// [1, 2); the class is exact unless a value is below 2^-480 times the
// largest one.
struct Expansion {
  std::vector<double> e;  // increasing magnitude, nonoverlapping
  void grow(double b) {
    std::vector<double> out;
    double q = b;
    for (double ei : e) {
      double s = q + ei;
      double bv = s - q;
      double av = s - bv;
      double err = (q - av) + (ei - bv);
      if (err != 0.0) out.push_back(err);
      q = s;
    }
    if (q != 0.0 || out.empty()) out.push_back(q);
    e.swap(out);
  }
  int sign() const {
    for (auto it = e.rbegin(); it != e.rend(); ++it) {
      if (*it > 0.0) return 1;
      if (*it < 0.0) return -1;
    }
    return 0;
  }
  double value() const {
    double s = 0.0;
    for (double v : e) s += v;
    return s;
  }
};

// The discriminant of (M, a, Q) scaled by 2^-k so that the largest
// magnitude lies in [1, 2): power-of-two scaling is exact, keeps every square
// finite and preserves the sign. *scale_out receives 2^k.
Expansion discriminant(double mass, double spin, double charge, double* scale_out = nullptr) {
  double big = std::fmax(std::fabs(mass), std::fmax(std::fabs(spin), std::fabs(charge)));
  int k = big > 0.0 ? std::ilogb(big) : 0;
  mass = std::ldexp(mass, -k);
  spin = std::ldexp(spin, -k);
  charge = std::ldexp(charge, -k);
  if (scale_out) *scale_out = std::ldexp(1.0, k);
  Expansion ex;
  auto add_square = [&ex](double v, double sign) {
    double hi = v * v;
    double lo = std::fma(v, v, -hi);
    ex.grow(sign * hi);
    ex.grow(sign * lo);
  };
  add_square(mass, 1.0);
  add_square(spin, -1.0);
  add_square(charge, -1.0);
  return ex;
}

// Binary decomposition helpers (numerics.rs:100-172): value = m * 2^e with
// m in [1, 2).
void decompose(double v, double* m, int* e) {
  int ex = 0;
  double fr = std::frexp(v, &ex);  // fr in [0.5, 1)
  *m = fr * 2.0;
  *e = ex - 1;
}

}  // namespace

// positive_product (numerics.rs:104-126): product of positive factors without
// intermediate overflow or underflow.
double positive_product(const double* f, int n) {
  for (int i = 0; i < n; ++i)
    if (f[i] == 0.0) return 0.0;
  double sig = 1.0;
  int ex = 0;
  for (int i = 0; i < n; ++i) {
    double m;
    int e;
    decompose(f[i], &m, &e);
    sig *= m;
    ex += e;
    if (sig >= 2.0) {
      sig *= 0.5;
      ex += 1;
    }
  }
  if (ex < -1075) return 0.0;
  return std::ldexp(sig, ex);
}

// normalized_quadratic_form_residual (numerics.rs:3-36).
double normalized_quadratic_form_residual(const double m[4][4], const double v[4]) {
  double mscale = 0.0, vscale = 0.0;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      if (!finite(m[r][c])) return NAN;
      mscale = std::fmax(mscale, std::fabs(m[r][c]));
    }
  for (int i = 0; i < 4; ++i) {
    if (!finite(v[i])) return NAN;
    vscale = std::fmax(vscale, std::fabs(v[i]));
  }
  if (mscale == 0.0 || vscale == 0.0) return 0.0;
  double sv[4];
  for (int i = 0; i < 4; ++i) sv[i] = v[i] / vscale;
  double contraction = 0.0, norm = 0.0;
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double term = (m[r][c] / mscale) * sv[r] * sv[c];
      contraction += term;
      norm += std::fabs(term);
    }
  if (norm == 0.0) return 0.0;
  double rel = std::fabs(contraction) / norm;
  double factors[4] = {norm, mscale, vscale, vscale};
  double den = std::fmin(positive_product(factors, 4), 1.0);
  return rel * den;
}

int compare_extremality(double mass, double spin, double charge) {
  int s = discriminant(mass, spin, charge).sign();
  return s > 0 ? kSubextremal : (s == 0 ? kExtremal : kSuperextremal);
}

// KerrNewmanSpacetime::validated_with_prefix + outer_horizon_radius
// (spacetime.rs:143-238).
int make_spacetime(double mass, double spin, double charge, double chart, Spacetime* out) {
  if (!finite(mass) || mass <= 0.0 || !finite(spin) || !finite(charge)) return 1;
  Spacetime st;
  st.mass = mass;
  st.spin = spin;
  st.charge = charge;
  st.chart = chart >= 0.0 ? 1.0 : -1.0;
  st.extremality = compare_extremality(mass, spin, charge);
  if (st.extremality == kSuperextremal) {
    st.has_horizon = false;
    st.outer_horizon = 0.0;
  } else if (st.extremality == kExtremal) {
    st.has_horizon = true;
    st.outer_horizon = mass;
  } else {
    st.has_horizon = true;
    double scale = 1.0;
    double d = discriminant(mass, spin, charge, &scale).value();
    st.outer_horizon = (std::ldexp(mass, -std::ilogb(scale)) + std::sqrt(d)) * scale;
  }
  if (st.has_horizon && !finite(st.outer_horizon)) return 1;
  *out = st;
  return 0;
}

namespace {

struct Scaled {
  double scale, x, y, z, spin, r2, r, sigma;
};

// scaled_coordinates (spacetime.rs:637-712).
int scaled_coordinates(const Spacetime& st, const double* p, Scaled* out) {
  double px = p[0], py = p[1], pz = p[2];
  if (!finite(px) || !finite(py) || !finite(pz)) return kGeomNonFinite;
  double scale = std::fmax(std::fmax(std::fmax(std::fabs(px), std::fabs(py)), std::fabs(pz)),
                           std::fabs(st.spin));
  if (scale == 0.0) return kGeomRingSingularity;
  Scaled s;
  s.scale = scale;
  s.x = px / scale;
  s.y = py / scale;
  s.z = pz / scale;
  s.spin = st.spin / scale;
  double a2 = s.spin * s.spin;
  if (px == 0.0 && py == 0.0 && pz != 0.0) {
    double r = std::fabs(s.z);
    double r2 = r * r;
    double sigma = r2 + a2;
    if (r > 0.0 && finite(sigma) && sigma > 0.0) {
      s.r2 = r2;
      s.r = r;
      s.sigma = sigma;
      *out = s;
      return kGeomOk;
    }
  }
  double rho2 = std::fma(s.x, s.x, std::fma(s.y, s.y, s.z * s.z));
  double b = rho2 - a2;
  double disc = std::fma(b, b, 4.0 * a2 * s.z * s.z);
  if (!finite(disc) || disc < 0.0) return kGeomInvalidDenominator;
  double root = std::sqrt(disc);
  double r2;
  if (b >= 0.0) {
    r2 = 0.5 * (b + root);
  } else {
    double den = root - b;
    if (den <= 0.0 || !finite(den)) return kGeomInvalidDenominator;
    r2 = 2.0 * a2 * s.z * s.z / den;
  }
  if (!finite(r2)) return kGeomInvalidDenominator;
  if (r2 <= 0.0) return b == 0.0 ? kGeomRingSingularity : kGeomChartBoundary;
  s.r2 = r2;
  s.r = std::sqrt(r2);
  s.sigma = r2 + a2 * s.z * s.z / r2;
  if (!finite(s.sigma) || s.sigma <= 0.0) return kGeomInvalidDenominator;
  *out = s;
  return kGeomOk;
}

int validate(const Geometry& g) {
  for (int i = 0; i < 3; ++i)
    if (!finite(g.radius_gradient[i]) || !finite(g.f_gradient[i])) return kGeomNonFinite;
  if (!finite(g.f)) return kGeomNonFinite;
  for (int i = 0; i < 4; ++i)
    if (!finite(g.null_covector[i])) return kGeomNonFinite;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j)
      if (!finite(g.null_vector_gradient[i][j])) return kGeomNonFinite;
  return kGeomOk;
}

// axis_geometry (spacetime.rs:584-635).
int axis_geometry(const Spacetime& st, const Scaled& s, Geometry* g) {
  double radius = s.scale * s.r;
  if (!finite(radius) || s.r <= 0.0 || s.sigma <= 0.0) return kGeomInvalidDenominator;
  double axis = s.z > 0.0 ? 1.0 : (s.z < 0.0 ? -1.0 : (std::signbit(s.z) ? -1.0 : 1.0));
  g->radius = radius;
  g->radius_gradient = {0.0, 0.0, axis};
  double m = st.mass / s.scale;
  double q = st.charge / s.scale;
  double num = std::fma(-q, q, 2.0 * m * s.r);
  g->f = num / s.sigma;
  V3 sig_grad = {0.0, 0.0, 2.0 * s.r * axis};
  V3 num_grad = {0.0, 0.0, axis * (2.0 * m)};
  for (int i = 0; i < 3; ++i)
    g->f_gradient[i] = (num_grad[i] * s.sigma - sig_grad[i] * num) / (s.sigma * s.sigma * s.scale);
  double bs = st.chart;
  g->null_covector = {1.0, 0.0, 0.0, bs * axis};
  g->null_vector = {-1.0, 0.0, 0.0, bs * axis};
  double ids = (1.0 / s.sigma) / s.scale;
  g->null_vector_gradient[0] = {0.0, bs * s.r * ids, -s.spin * ids, 0.0};
  g->null_vector_gradient[1] = {0.0, s.spin * ids, bs * s.r * ids, 0.0};
  g->null_vector_gradient[2] = {0.0, 0.0, 0.0, 0.0};
  return validate(*g);
}

}  // namespace

// geometry (spacetime.rs:492-582).
int geometry(const Spacetime& st, const double* xyz, Geometry* g) {
  Scaled s;
  int err = scaled_coordinates(st, xyz, &s);
  if (err) return err;
  if (s.x == 0.0 && s.y == 0.0) return axis_geometry(st, s, g);
  double radius = s.scale * s.r;
  if (!finite(radius)) return kGeomInvalidDenominator;
  double a2 = s.spin * s.spin;
  V3 rg = {s.x * s.r / s.sigma, s.y * s.r / s.sigma, s.z * (s.r2 + a2) / (s.r * s.sigma)};
  double m = st.mass / s.scale;
  double q = st.charge / s.scale;
  double num = std::fma(-q, q, 2.0 * m * s.r);
  double f = num / s.sigma;
  double vert = a2 * s.z * s.z / s.r2;
  double srf = 2.0 * (s.r - vert / s.r);
  V3 sg = {srf * rg[0], srf * rg[1], std::fma(srf, rg[2], 2.0 * a2 * s.z / s.r2)};
  V3 ng = {rg[0] * (2.0 * m), rg[1] * (2.0 * m), rg[2] * (2.0 * m)};
  V3 fg;
  for (int i = 0; i < 3; ++i) fg[i] = (ng[i] * s.sigma - sg[i] * num) / (s.sigma * s.sigma * s.scale);
  double K = s.r2 + a2;
  if (K <= 0.0 || !finite(K)) return kGeomInvalidDenominator;
  double bs = st.chart;
  double cs = bs * s.spin;
  V3 ell = {bs * (std::fma(cs, s.y, s.r * s.x) / K), bs * (std::fma(cs, -s.x, s.r * s.y) / K),
            bs * (s.z / s.r)};
  g->radius = radius;
  g->radius_gradient = rg;
  g->f = f;
  g->f_gradient = fg;
  g->null_covector = {1.0, ell[0], ell[1], ell[2]};
  g->null_vector = {-1.0, ell[0], ell[1], ell[2]};
  double coords[3] = {s.x, s.y, s.z};
  for (int i = 0; i < 3; ++i) {
    double ri = rg[i];
    double dx = i == 0 ? 1.0 : 0.0, dy = i == 1 ? 1.0 : 0.0, dz = i == 2 ? 1.0 : 0.0;
    double nx = std::fma(cs, dy, std::fma(ri, s.x, s.r * dx));
    double ny = std::fma(cs, -dx, std::fma(ri, s.y, s.r * dy));
    double rd = 2.0 * s.r * ri;
    double ddx = std::fma(-bs * ell[0] * rd, 1.0 / K, nx / K) / s.scale;
    double ddy = std::fma(-bs * ell[1] * rd, 1.0 / K, ny / K) / s.scale;
    double ddz = (dz / s.r - coords[2] * ri / s.r2) / s.scale;
    g->null_vector_gradient[i] = {0.0, bs * ddx, bs * ddy, bs * ddz};
  }
  return validate(*g);
}

// metric_covariant / metric_inverse (spacetime.rs:738-764).
void metric_covariant(const Geometry& g, double out[4][4]) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double mk = r == c ? (r == 0 ? -1.0 : 1.0) : 0.0;
      out[r][c] = std::fma(g.f * g.null_covector[r], g.null_covector[c], mk);
    }
}

void metric_inverse(const Geometry& g, double out[4][4]) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double mk = r == c ? (r == 0 ? -1.0 : 1.0) : 0.0;
      out[r][c] = std::fma(-g.f * g.null_vector[r], g.null_vector[c], mk);
    }
}

namespace {
double dot4(const V4& a, const double* b) {
  double s = 0.0;
  for (int i = 0; i < 4; ++i) s += a[i] * b[i];
  return s;
}
}  // namespace

// hamiltonian_rhs (spacetime.rs:305-338): full 8-component Hamilton system
// with the complete null-vector Jacobian.
int hamiltonian_rhs(const Spacetime& st, const S8& y, S8* out) {
  for (double v : y)
    if (!finite(v)) return kGeomNonFinite;
  Geometry g;
  int err = geometry(st, &y[1], &g);
  if (err) return err;
  const double* p = &y[4];
  double contraction = dot4(g.null_vector, p);
  double raised[4] = {-p[0], p[1], p[2], p[3]};
  S8 d;
  for (int i = 0; i < 4; ++i) d[i] = std::fma(g.f * contraction, -g.null_vector[i], raised[i]);
  d[4] = 0.0;
  for (int i = 0; i < 3; ++i) {
    double ndc = 0.0;
    for (int k = 0; k < 4; ++k) ndc += p[k] * g.null_vector_gradient[i][k];
    d[5 + i] = std::fma(g.f * contraction, ndc, 0.5 * contraction * contraction * g.f_gradient[i]);
  }
  for (double v : d)
    if (!finite(v)) return kGeomNonFinite;
  *out = d;
  return kGeomOk;
}

// invariants + normalized_null_residual (spacetime.rs:345-401).
int invariants(const Spacetime& st, const S8& y, Invariants* out) {
  Geometry g;
  int err = geometry(st, &y[1], &g);
  if (err) return err;
  double radius = g.radius;
  const double* p = &y[4];
  double x = y[1], yy = y[2], z = y[3];
  double energy = -p[0];
  double lz = std::fma(yy, -p[1], x * p[2]);
  double rho = std::hypot(x, yy);
  double a2 = st.spin * st.spin;
  double carter;
  if (rho == 0.0) {
    double tp2 = std::fma(p[1], p[1], p[2] * p[2]);
    double lead = std::fma(radius, radius, a2) * tp2;
    carter = std::fma(-a2 * energy, energy, lead);
  } else {
    double cth = z / radius;
    double sth = rho / std::sqrt(std::fma(radius, radius, a2));
    double proj = std::fma(x, p[1], yy * p[2]);
    double ptheta = std::fma(-radius * sth, p[3], (cth / sth) * proj);
    double lzs = lz / sth;
    double ang = std::fma(-a2 * energy, energy, lzs * lzs);
    carter = std::fma(cth * cth, ang, ptheta * ptheta);
  }
  double inv[4][4];
  metric_inverse(g, inv);
  double nr = normalized_quadratic_form_residual(inv, p);
  if (!finite(nr)) return kGeomNonFinite;
  if (!finite(energy) || !finite(lz) || !finite(carter)) return kGeomNonFinite;
  *out = {energy, lz, carter, nr};
  return kGeomOk;
}

// radial_velocity (spacetime.rs:408-414).
int radial_velocity(const Spacetime& st, const S8& y, double* out) {
  Geometry g;
  int err = geometry(st, &y[1], &g);
  if (err) return err;
  S8 d;
  err = hamiltonian_rhs(st, y, &d);
  if (err) return err;
  *out = g.radius_gradient[0] * d[1] + g.radius_gradient[1] * d[2] + g.radius_gradient[2] * d[3];
  return kGeomOk;
}

// singularity_measure (spacetime.rs:114-126, 421-427): r^4 + a^2 z^2.
int singularity_measure(const Spacetime& st, const double* xyz, double* out) {
  Scaled s;
  int err = scaled_coordinates(st, xyz, &s);
  if (err) return err;
  double a2 = s.spin * s.spin;
  double f[5] = {s.scale, s.scale, s.scale, s.scale, std::fma(s.r2, s.r2, a2 * s.z * s.z)};
  double m = positive_product(f, 5);
  if (!finite(m)) return kGeomNonFinite;
  *out = m;
  return kGeomOk;
}

int singularity_guard_residual(const Spacetime& st, const double* xyz, double guard, double* out) {
  if (!finite(guard) || guard <= 0.0) return kGeomInvalidDenominator;
  Scaled s;
  int err = scaled_coordinates(st, xyz, &s);
  if (err) return err;
  double a2 = s.spin * s.spin;
  double f[5] = {s.scale, s.scale, s.scale, s.scale, std::fma(s.r2, s.r2, a2 * s.z * s.z)};
  double m = positive_product(f, 5);
  *out = std::isinf(m) ? std::numeric_limits<double>::max() : m - guard;
  return kGeomOk;
}

}  // namespace orc
