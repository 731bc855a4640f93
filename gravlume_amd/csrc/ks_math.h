// Reduced Cartesian Kerr-Schild null-geodesic right-hand side.
//
// The state is the spatial position x and covariant spatial momentum p with
// the per-ray constant E = -p_t (stationarity). With f the Kerr-Schild
// profile, l the spatial principal null vector and q = E + l.p, the reduced
// Hamilton system is
//     x' = p - f q l,   p' = f q w + q^2/2 grad f,   t' = E + f q,
// where w = grad(l.p) at fixed p = e + C grad r (docs/physics.md:251-280 of the
// reference; the reference's WGSL form is kerr_schild_dynamics.wgsl:18-269,
// its f64 CPU form spacetime.rs:305-338).
//
// Everything here is a template over the scalar type so that one source
// serves the f32 hot path, the f64 near-horizon / root-solve islands and a
// host-side operation-counting type that fixes the flop constants used for
// the roofline.
#pragma once

#if defined(__HIP__)
#define GL_HD __attribute__((host, device)) inline __attribute__((always_inline))
#else
#define GL_HD inline
#endif

namespace gl {

template <class T>
struct Geom {
  T r;           // oblate radius
  T f;           // Kerr-Schild profile (2 M r - Q^2) r^2 / (r^4 + a^2 z^2)
  T lx, ly, lz;  // chart-signed spatial principal null vector
  T grx, gry, grz;  // grad r
  T gfx, gfy, gfz;  // grad f
  T kinv, rinv;     // 1/(r^2+a^2), 1/r
  bool ok;
};

template <class T>
GL_HD T gl_abs(T v) {
  return v < T(0) ? -v : v;
}
// |v| as a source modifier of the consuming instruction (no instruction).
template <>
GL_HD float gl_abs<float>(float v) {
  return __builtin_fabsf(v);
}
template <>
GL_HD double gl_abs<double>(double v) {
  return __builtin_fabs(v);
}

// Square root, reciprocal and reciprocal square root. On the f32 device
// path these are the hardware v_sqrt_f32 / v_rcp_f32 / v_rsq_f32 (about one
// ulp) instead of the multi-instruction IEEE sequences; the f64 path and the
// host keep IEEE operations.
template <class T>
GL_HD T gl_sqrt(T v);
template <>
GL_HD float gl_sqrt<float>(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(v);
#else
  return __builtin_sqrtf(v);
#endif
}
template <>
GL_HD double gl_sqrt<double>(double v) {
  return __builtin_sqrt(v);
}

template <class T>
GL_HD T gl_rcp(T v) {
  return T(1) / v;
}
template <>
GL_HD float gl_rcp<float>(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(v);
#else
  return 1.0f / v;
#endif
}

// f64 on the device: the hardware v_rcp_f64 / v_rsq_f64 estimate refined by
// two Newton steps (a few ulp, for finite nonzero v) instead of the IEEE
// division and square-root sequences.
template <>
GL_HD double gl_rcp<double>(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(v);
  r = __builtin_fma(__builtin_fma(-v, r, 1.0), r, r);
  r = __builtin_fma(__builtin_fma(-v, r, 1.0), r, r);
  return r;
#else
  return 1.0 / v;
#endif
}

template <class T>
GL_HD T gl_rsqrt(T v) {
  return T(1) / gl_sqrt<T>(v);
}
template <>
GL_HD double gl_rsqrt<double>(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(v);
  const double hv = 0.5 * v;
  y = y * __builtin_fma(-hv * y, y, 1.5);
  y = y * __builtin_fma(-hv * y, y, 1.5);
  return y;
#else
  return 1.0 / __builtin_sqrt(v);
#endif
}
template <>
GL_HD float gl_rsqrt<float>(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(v);
#else
  return 1.0f / __builtin_sqrtf(v);
#endif
}

// Oblate radius and the metric quantities at a point. mass, spin, spin^2,
// charge^2 and the chart sign s (+1 ingoing, -1 outgoing) are scene
// constants. The radius is the non-negative root of
// r^4 - (rho^2 - a^2) r^2 - a^2 z^2 = 0, taken in the conjugate form when
// rho^2 < a^2 to avoid cancellation (spacetime.rs:675-698).
template <class T>
GL_HD Geom<T> geometry(T x, T y, T z, T mass, T spin, T spin2, T charge2, T s) {
  Geom<T> g;
  T rho2 = x * x + y * y + z * z;
  T b = rho2 - spin2;
  T az2 = spin2 * (z * z);
  T disc = b * b + T(4) * az2;
  T sinv = gl_rsqrt<T>(disc);
  T S = disc * sinv;  // (r^4 + a^2 z^2) / r^2
  T r2 = T(0.5) * (b + S);
  if (__builtin_expect(b < T(0), 0)) r2 = (T(2) * az2) * gl_rcp<T>(S - b);
  T rinv = gl_rsqrt<T>(r2);
  T r = r2 * rinv;
  T kinv = gl_rcp<T>(r2 + spin2);
  T num = T(2) * mass * r - charge2;
  T f = num * sinv;
  T rs = r * sinv;
  g.grx = x * rs;
  g.gry = y * rs;
  g.grz = z * (r2 + spin2) * rinv * sinv;
  // grad S = 2 (b x, b y, z (rho^2 + a^2)) / S
  T two_sinv = T(2) * sinv;
  T gsx = b * x * two_sinv;
  T gsy = b * y * two_sinv;
  T gsz = z * (rho2 + spin2) * two_sinv;
  T tm = T(2) * mass;
  g.gfx = (tm * g.grx - f * gsx) * sinv;
  g.gfy = (tm * g.gry - f * gsy) * sinv;
  g.gfz = (tm * g.grz - f * gsz) * sinv;
  T sr = s * r;
  g.lx = (sr * x + spin * y) * kinv;
  g.ly = (sr * y - spin * x) * kinv;
  g.lz = s * z * rinv;
  g.r = r;
  g.f = f;
  g.kinv = kinv;
  g.rinv = rinv;
  g.ok = (r2 > T(0)) && (S > T(0)) && (rho2 < T(1e30));
  return g;
}

template <class T>
struct Deriv {
  T dt;            // coordinate-time rate t'
  T dx, dy, dz;    // x'
  T dpx, dpy, dpz; // p'
};

// Right-hand side from a precomputed geometry (the contraction w is formed
// without the full 3x3 Jacobian of l).
template <class T>
GL_HD Deriv<T> rhs_from_geometry(const Geom<T>& g, T x, T y, T z, T px, T py, T pz, T E,
                                 T spin, T s) {
  Deriv<T> d;
  T lp = g.lx * px + g.ly * py + g.lz * pz;
  T q = E + lp;
  T fq = g.f * q;
  d.dx = px - fq * g.lx;
  d.dy = py - fq * g.ly;
  d.dz = pz - fq * g.lz;
  d.dt = E + fq;
  T sr = s * g.r;
  T ex = (sr * px - spin * py) * g.kinv;
  T ey = (spin * px + sr * py) * g.kinv;
  T ez = s * pz * g.rinv;
  T c = (s * (x * px + y * py) - T(2) * g.r * (g.lx * px + g.ly * py)) * g.kinv -
        s * z * pz * (g.rinv * g.rinv);
  T wx = ex + c * g.grx;
  T wy = ey + c * g.gry;
  T wz = ez + c * g.grz;
  T hq2 = T(0.5) * q * q;
  d.dpx = fq * wx + hq2 * g.gfx;
  d.dpy = fq * wy + hq2 * g.gfy;
  d.dpz = fq * wz + hq2 * g.gfz;
  return d;
}

// The right-hand side alone, for intermediate DP5 stages that need neither
// the geometry nor its gradients afterwards. Same equations as geometry() +
// rhs_from_geometry(), regrouped: with grad r = (x rs, y rs, z G) and
// grad f = (x Cxy, y Cxy, z Cz),
//     p' = f q e + (x A, y A, z Az),  A = f q c rs + q^2/2 Cxy,
//                                     Az = f q c G + q^2/2 Cz,
// so the gradient vectors are never formed. Returns the geometry's validity.
template <class T>
GL_HD bool rhs_only(T x, T y, T z, T px, T py, T pz, T E, T mass, T spin, T spin2, T charge2,
                    T s, Deriv<T>& d) {
  const T rho2 = x * x + y * y + z * z;
  const T b = rho2 - spin2;
  const T az2 = spin2 * (z * z);
  const T disc = b * b + T(4) * az2;
  const T sinv = gl_rsqrt<T>(disc);
  const T S = disc * sinv;
  T r2 = T(0.5) * (b + S);
  if (__builtin_expect(b < T(0), 0)) r2 = (T(2) * az2) * gl_rcp<T>(S - b);
  const T rinv = gl_rsqrt<T>(r2);
  const T r = r2 * rinv;
  const T D = r2 + spin2;
  const T kinv = gl_rcp<T>(D);
  const T tm = T(2) * mass;
  const T f = (tm * r - charge2) * sinv;
  const T rs = r * sinv;
  const T G = D * rinv * sinv;
  const T tfs = T(2) * f * sinv;
  const T cxy = (tm * rs - tfs * b) * sinv;
  const T cz = (tm * G - tfs * (rho2 + spin2)) * sinv;
  const T sr = s * r;
  const T lx = (sr * x + spin * y) * kinv;
  const T ly = (sr * y - spin * x) * kinv;
  const T lz = s * z * rinv;
  const T q = E + (lx * px + ly * py + lz * pz);
  const T fq = f * q;
  d.dx = px - fq * lx;
  d.dy = py - fq * ly;
  d.dz = pz - fq * lz;
  d.dt = E + fq;
  const T ex = (sr * px - spin * py) * kinv;
  const T ey = (spin * px + sr * py) * kinv;
  const T ez = s * pz * rinv;
  const T c = (s * (x * px + y * py) - T(2) * r * (lx * px + ly * py)) * kinv -
              s * z * pz * (rinv * rinv);
  const T hq2 = T(0.5) * q * q;
  const T fqc = fq * c;
  const T axy = fqc * rs + hq2 * cxy;
  const T az = fqc * G + hq2 * cz;
  d.dpx = fq * ex + x * axy;
  d.dpy = fq * ey + y * axy;
  d.dpz = fq * ez + z * az;
  return (r2 > T(0)) && (S > T(0)) && (rho2 < T(1e30));
}

// Coordinate radial rate dr/dlambda along the Hamilton flow.
template <class T>
GL_HD T radial_rate(const Geom<T>& g, const Deriv<T>& d) {
  return g.grx * d.dx + g.gry * d.dy + g.grz * d.dz;
}

// Null residual |p.x'| normalised by its term norm (the reference GPU form,
// geodesic_events.wgsl:381-385): with p_t = -E the contraction is
// -E t' + p.x'.
template <class T>
GL_HD T null_residual(const Deriv<T>& d, T px, T py, T pz, T E) {
  T c = -E * d.dt + px * d.dx + py * d.dy + pz * d.dz;
  T n = gl_abs<T>(E * d.dt) + gl_abs<T>(px * d.dx) + gl_abs<T>(py * d.dy) +
        gl_abs<T>(pz * d.dz);
  T den = n > T(1) ? n : T(1);
  return gl_abs<T>(c) * gl_rcp<T>(den);
}

// Carter constant in the axis-regular Cartesian form (docs/physics.md:318-331
// of the reference): with D = r^2 + a^2, S = x px + y py,
//   Q = z^2/r^2 (D (px^2+py^2) - a^2 E^2) - 2 z S pz + r^2 (x^2+y^2)/D pz^2.
template <class T>
GL_HD T carter(const Geom<T>& g, T x, T y, T z, T px, T py, T pz, T E, T spin2) {
  T D = g.r * g.r + spin2;
  T cz = z * g.rinv;
  T S = x * px + y * py;
  return cz * cz * (D * (px * px + py * py) - spin2 * E * E) - T(2) * z * S * pz +
         g.r * g.r * (x * x + y * y) * g.kinv * pz * pz;
}

}  // namespace gl
