// Interval radial-capture certificate (the reference's conservative capture
// test, geodesic_acceleration.wgsl:34-279; docs/research/
// gpu-geodesic-acceleration.md:182-189).
//
// With the constants of motion E, L_z and Carter Q of a backward ray and
// K = (L - a E)^2 + Q, the Kerr-Newman radial potential is
//     R(r) = E^2 r^4 + (-2 a E (L - a E) - K) r^2 + 2 K r - a^2 Q - q^2 K .
// If R > 0 on [r_+, r_obs] the radial motion has no turning point between the
// horizon and the observer, so a ray that starts with r decreasing along the
// backward traversal reaches the horizon: a capture, proven without
// integrating. Positivity is shown on 12 equal segments by the degree-4
// Bernstein coefficients of R on each segment, all evaluated in f32 interval
// arithmetic in which every operation's bounds are moved one ulp outward
// (round-to-nearest errs by at most half an ulp, so each true result stays
// inside). The inputs E, L, Q and r_obs are widened by 2^-12 (1 + |v|), a and
// q by one ulp. Any doubt (outside the domain |a| <= 0.9, a^2 + q^2 <
// 1 - 2^-12, r_obs <= 256, off-axis by r_obs/16, a clearly inward backward
// start, or a non-positive coefficient) means "not certified", and the ray is
// traced in full.
#pragma once

#include <cstdint>
#include <cstring>

#include "ks_math.h"

namespace gl {

struct Interval {
  float lo, hi;
};

GL_HD uint32_t cert_bits(float v) {
  uint32_t b;
  __builtin_memcpy(&b, &v, 4);
  return b;
}
GL_HD float cert_float(uint32_t b) {
  float v;
  __builtin_memcpy(&v, &b, 4);
  return v;
}

// One ulp below / above v; across zero and for subnormals the smallest
// normal magnitude (a conservative step).
GL_HD float cert_down(float v) {
  const float tiny = 1.17549435e-38f;  // 2^-126
  if (gl_abs<float>(v) < tiny) return -tiny;
  const uint32_t b = cert_bits(v);
  return cert_float(v > 0.0f ? b - 1u : b + 1u);
}
GL_HD float cert_up(float v) {
  const float tiny = 1.17549435e-38f;
  if (gl_abs<float>(v) < tiny) return tiny;
  const uint32_t b = cert_bits(v);
  return cert_float(v > 0.0f ? b + 1u : b - 1u);
}

GL_HD float gl_min_f(float a, float b) { return a < b ? a : b; }
GL_HD float gl_max_f(float a, float b) { return a > b ? a : b; }

GL_HD Interval iv_point(float v) { return Interval{v, v}; }
GL_HD Interval iv_ulp(float v) { return Interval{cert_down(v), cert_up(v)}; }
GL_HD Interval iv_add(Interval a, Interval b) { return Interval{cert_down(a.lo + b.lo), cert_up(a.hi + b.hi)}; }
GL_HD Interval iv_sub(Interval a, Interval b) { return Interval{cert_down(a.lo - b.hi), cert_up(a.hi - b.lo)}; }
GL_HD Interval iv_neg(Interval a) { return Interval{-a.hi, -a.lo}; }
GL_HD Interval iv_mul(Interval a, Interval b) {
  const float p0 = a.lo * b.lo, p1 = a.lo * b.hi, p2 = a.hi * b.lo, p3 = a.hi * b.hi;
  const float lo = gl_min_f(gl_min_f(p0, p1), gl_min_f(p2, p3));
  const float hi = gl_max_f(gl_max_f(p0, p1), gl_max_f(p2, p3));
  return Interval{cert_down(lo), cert_up(hi)};
}
GL_HD Interval iv_sqr(Interval a) {
  const float s0 = a.lo * a.lo, s1 = a.hi * a.hi;
  if (a.lo <= 0.0f && a.hi >= 0.0f) return Interval{0.0f, cert_up(gl_max_f(s0, s1))};
  return Interval{cert_down(gl_min_f(s0, s1)), cert_up(gl_max_f(s0, s1))};
}
GL_HD Interval iv_scale(Interval a, float s) { return iv_mul(a, iv_point(s)); }
GL_HD bool iv_finite(Interval a) {
  return a.lo == a.lo && a.hi == a.hi && gl_abs<float>(a.lo) <= 3.4028235e38f &&
         gl_abs<float>(a.hi) <= 3.4028235e38f && a.lo <= a.hi;
}
// v widened by 2^-12 (1 + |v|).
GL_HD Interval iv_input(float v) {
  const float env = 1.0f / 4096.0f;
  const float pad = iv_mul(iv_add(iv_point(gl_abs<float>(v)), iv_point(1.0f)), iv_point(env)).hi;
  return Interval{cert_down(v - pad), cert_up(v + pad)};
}

// R(lower + width t) > 0 for t in [0, 1] by its Bernstein coefficients.
GL_HD bool cert_segment_positive(Interval lo, Interval w, Interval c4, Interval c2, Interval c1,
                                 Interval c0) {
  const Interval lo2 = iv_sqr(lo), lo3 = iv_mul(lo2, lo), lo4 = iv_sqr(lo2);
  const Interval w2 = iv_sqr(w), w3 = iv_mul(w2, w), w4 = iv_sqr(w2);
  // power-basis coefficients of R(lo + w t)
  const Interval p0 = iv_add(iv_add(iv_mul(c4, lo4), iv_mul(c2, lo2)), iv_add(iv_mul(c1, lo), c0));
  const Interval p1 =
      iv_mul(w, iv_add(iv_add(iv_scale(iv_mul(c4, lo3), 4.0f), iv_scale(iv_mul(c2, lo), 2.0f)), c1));
  const Interval p2 = iv_mul(w2, iv_add(iv_scale(iv_mul(c4, lo2), 6.0f), c2));
  const Interval p3 = iv_scale(iv_mul(iv_mul(c4, lo), w3), 4.0f);
  const Interval p4 = iv_mul(c4, w4);
  const Interval sixth = Interval{cert_down(1.0f / 6.0f), cert_up(1.0f / 6.0f)};
  // degree-4 Bernstein coefficients
  const Interval b0 = p0;
  const Interval b1 = iv_add(p0, iv_scale(p1, 0.25f));
  const Interval b2 = iv_add(iv_add(p0, iv_scale(p1, 0.5f)), iv_mul(p2, sixth));
  const Interval b3 = iv_add(iv_add(p0, iv_scale(p1, 0.75f)), iv_add(iv_scale(p2, 0.5f), iv_scale(p3, 0.25f)));
  const Interval b4 = iv_add(iv_add(p0, p1), iv_add(iv_add(p2, p3), p4));
  const Interval bs[5] = {b0, b1, b2, b3, b4};
  for (int i = 0; i < 5; ++i)
    if (!iv_finite(bs[i]) || !(bs[i].lo > 0.0f)) return false;
  return true;
}

// The certificate. mass must be 1; r_plus is the outer horizon, r_obs the
// observer's oblate radius, rho2 its x^2 + y^2, dr_future the radial rate
// dr/dlambda of the future-directed momentum (positive: the backward ray
// heads inward).
GL_HD bool capture_certified(float mass, float spin, float charge, float r_plus, float r_obs,
                             float rho2, float dr_future, float E, float L, float Q) {
  const float env = 1.0f / 4096.0f;
  const float ext = spin * spin + charge * charge;
  const float margin = 0.0625f * r_obs;
  if (!(mass == 1.0f) || !(gl_abs<float>(spin) <= 0.9f) || !(ext == ext) ||
      !(ext < 1.0f - env) || !(r_plus > 0.0f) || !(r_obs > r_plus) || !(r_obs <= 256.0f) ||
      !(rho2 > margin * margin) || !(dr_future > env) || !(gl_abs<float>(E) <= 4096.0f) ||
      !(gl_abs<float>(L) <= 4096.0f) || !(gl_abs<float>(Q) <= 4096.0f))
    return false;
  const Interval e = iv_input(E), l = iv_input(L), q = iv_input(Q);
  const Interval a = iv_ulp(spin), ch = iv_ulp(charge);
  const Interval shifted = iv_sub(l, iv_mul(a, e));
  const Interval K = iv_add(iv_sqr(shifted), q);
  if (!iv_finite(K) || !(K.lo > 0.0f)) return false;
  const Interval c4 = iv_sqr(e);
  const Interval c2 = iv_sub(iv_neg(iv_scale(iv_mul(iv_mul(e, a), shifted), 2.0f)), K);
  const Interval c1 = iv_scale(K, 2.0f);
  const Interval c0 = iv_sub(iv_neg(iv_mul(iv_sqr(a), q)), iv_mul(iv_sqr(ch), K));
  const Interval hor = iv_ulp(r_plus);
  const Interval robs = iv_input(r_obs);
  const Interval twelfth = Interval{cert_down(1.0f / 12.0f), cert_up(1.0f / 12.0f)};
  const Interval w = iv_mul(iv_sub(robs, hor), twelfth);
  if (!iv_finite(w) || !(w.lo > 0.0f)) return false;
  for (int s = 0; s < 12; ++s) {
    const Interval lo = iv_add(hor, iv_scale(w, float(s)));
    if (!cert_segment_positive(lo, w, c4, c2, c1, c0)) return false;
  }
  return true;
}

}  // namespace gl
