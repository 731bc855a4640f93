// Per-ray adaptive Dormand-Prince 5(4) trace of a Kerr-Newman null geodesic:
// one ray per lane, state in registers.
//
// Semantics follow the reference's trace state machine (CPU oracle
// crates/gravlume-reference/src/geodesic.rs:129-212, GPU protocol
// crates/gravlume-render/src/shaders/geodesic_integration.wgsl:22-408):
//   * events: singularity guard, horizon capture, equatorial surface,
//     escape, each armed once its event function leaves the arming band
//     (geodesic.rs:531-573), localized on the dense output of the accepted
//     step and ordered by affine fraction with a tie tolerance
//     (geodesic.rs:722-740);
//   * branch key: initial polar side, radial turnings and equatorial
//     crossings committed strictly before the terminal fraction, signed
//     azimuth winding (geodesic.rs:330-418);
//   * compensated (Kahan) accumulation of position, coordinate time and
//     affine parameter; step controller of policy.rs:137-173.
// The RHS and the DP5 stages run in the template scalar T: float on the hot
// path, double in the near-horizon finisher.
#pragma once

#include <cmath>

#include "gl_types.h"
#include "ks_math.h"

namespace gl {

// Dormand-Prince 5(4) tableau (Dormand & Prince 1980).
#define GL_A21 (1.0 / 5.0)
#define GL_A31 (3.0 / 40.0)
#define GL_A32 (9.0 / 40.0)
#define GL_A41 (44.0 / 45.0)
#define GL_A42 (-56.0 / 15.0)
#define GL_A43 (32.0 / 9.0)
#define GL_A51 (19372.0 / 6561.0)
#define GL_A52 (-25360.0 / 2187.0)
#define GL_A53 (64448.0 / 6561.0)
#define GL_A54 (-212.0 / 729.0)
#define GL_A61 (9017.0 / 3168.0)
#define GL_A62 (-355.0 / 33.0)
#define GL_A63 (46732.0 / 5247.0)
#define GL_A64 (49.0 / 176.0)
#define GL_A65 (-5103.0 / 18656.0)
#define GL_B1 (35.0 / 384.0)
#define GL_B3 (500.0 / 1113.0)
#define GL_B4 (125.0 / 192.0)
#define GL_B5 (-2187.0 / 6784.0)
#define GL_B6 (11.0 / 84.0)
// error weights b5 - b4
#define GL_E1 (35.0 / 384.0 - 5179.0 / 57600.0)
#define GL_E3 (500.0 / 1113.0 - 7571.0 / 16695.0)
#define GL_E4 (125.0 / 192.0 - 393.0 / 640.0)
#define GL_E5 (-2187.0 / 6784.0 + 92097.0 / 339200.0)
#define GL_E6 (11.0 / 84.0 - 187.0 / 2100.0)
#define GL_E7 (-1.0 / 40.0)

// Quartic continuous extension of DP5 (Shampine 1986): weight of stage i at
// fraction theta is theta * (d0 + theta * (d1 + theta * (d2 + theta * d3))).
// dense_table holds the d coefficients (the same numbers dense_weights and
// z_root spell out inline, where the f32 trace keeps them as literals).
template <class T>
GL_HD void dense_table(T d[7][4]) {
  const T t[7][4] = {
      {T(1.0), T(-8048581381.0 / 2820520608.0), T(8663915743.0 / 2820520608.0),
       T(-12715105075.0 / 11282082432.0)},
      {T(0), T(0), T(0), T(0)},
      {T(0), T(131558114200.0 / 32700410799.0), T(-68118460800.0 / 10900136933.0),
       T(87487479700.0 / 32700410799.0)},
      {T(0), T(-1754552775.0 / 470086768.0), T(14199869525.0 / 1410260304.0),
       T(-10690763975.0 / 1880347072.0)},
      {T(0), T(127303824393.0 / 49829197408.0), T(-318862633887.0 / 49829197408.0),
       T(701980252875.0 / 199316789632.0)},
      {T(0), T(-282668133.0 / 205662961.0), T(2019193451.0 / 616988883.0),
       T(-1453857185.0 / 822651844.0)},
      {T(0), T(40617522.0 / 29380423.0), T(-110615467.0 / 29380423.0),
       T(69997945.0 / 29380423.0)}};
#pragma unroll
  for (int i = 0; i < 7; ++i)
#pragma unroll
    for (int k = 0; k < 4; ++k) d[i][k] = t[i][k];
}

template <class T>
GL_HD void dense_weights(T th, T w[7]) {
  const T d[7][4] = {
      {T(1.0), T(-8048581381.0 / 2820520608.0), T(8663915743.0 / 2820520608.0),
       T(-12715105075.0 / 11282082432.0)},
      {T(0), T(0), T(0), T(0)},
      {T(0), T(131558114200.0 / 32700410799.0), T(-68118460800.0 / 10900136933.0),
       T(87487479700.0 / 32700410799.0)},
      {T(0), T(-1754552775.0 / 470086768.0), T(14199869525.0 / 1410260304.0),
       T(-10690763975.0 / 1880347072.0)},
      {T(0), T(127303824393.0 / 49829197408.0), T(-318862633887.0 / 49829197408.0),
       T(701980252875.0 / 199316789632.0)},
      {T(0), T(-282668133.0 / 205662961.0), T(2019193451.0 / 616988883.0),
       T(-1453857185.0 / 822651844.0)},
      {T(0), T(40617522.0 / 29380423.0), T(-110615467.0 / 29380423.0),
       T(69997945.0 / 29380423.0)}};
#pragma unroll
  for (int i = 0; i < 7; ++i) w[i] = th * (d[i][0] + th * (d[i][1] + th * (d[i][2] + th * d[i][3])));
}

// d w_i / d theta.
template <class T>
GL_HD void dense_slopes(T th, T w[7]) {
  T d[7][4];
  dense_table<T>(d);
#pragma unroll
  for (int i = 0; i < 7; ++i)
    w[i] = d[i][0] + th * (T(2) * d[i][1] + th * (T(3) * d[i][2] + th * (T(4) * d[i][3])));
}

// Without a horizon, rays that come within this oblate radius of the ring
// singularity finish in f64: binary32 cannot resolve the approach to the
// singularity guard (r^4 + a^2 z^2 = 2^-40), where the oblate radius comes
// from the cancellation x^2 + y^2 + z^2 - a^2, and its step control
// underflows before the guard is reached.
constexpr float kRingDeferRadius = 0.25f;

// Scene constants in the working precision.
template <class T>
struct SceneT {
  T mass, spin, spin2, charge2, chart;
  // r_defer: f32 rays below it finish in f64 (1.05 r_+, or for a spacetime
  // without a horizon kRingDeferRadius around the ring singularity)
  T r_capture, r_defer, r_escape, disk_in, disk_out;
  T band, rtol, atol, h_min, h_max, h_max_per_r, tie_tol, drift_limit, guard;
  T safety, min_factor, max_factor;  // step controller (policy.rs:160-173)
  uint32_t max_accepted, max_rejects, options, reproject_interval;
};

// The packed scene's mass is exactly 1 (the plan rejects any other), so the
// mass is the literal 1 here and its products fold away in the kernels.
GL_HD SceneT<float> scene_f32(const SceneDev& s) {
  SceneT<float> c;
  (void)s.mass;
  c.mass = 1.0f;
  c.spin = s.spin;
  c.spin2 = s.spin2;
  c.charge2 = s.charge2;
  c.chart = s.chart;
  c.r_capture = s.r_capture;
  c.r_defer = s.r_horizon > 0.0f ? 1.05f * s.r_horizon : kRingDeferRadius;
  c.r_escape = s.r_escape;
  c.disk_in = s.disk_in;
  c.disk_out = s.disk_out;
  c.band = s.arming_band;
  c.rtol = s.rtol;
  c.atol = s.atol;
  c.h_min = s.h_min;
  c.h_max = s.h_max;
  c.h_max_per_r = s.h_max_per_r;
  c.tie_tol = s.tie_tol;
  c.drift_limit = s.drift_limit;
  c.guard = s.guard;
  c.safety = s.safety;
  c.min_factor = s.min_factor;
  c.max_factor = s.max_factor;
  c.max_accepted = s.max_accepted;
  c.max_rejects = s.max_rejects;
  c.options = s.options;
  c.reproject_interval = s.reproject_interval;
  return c;
}

GL_HD SceneT<double> scene_f64(const SceneDev& s) {
  SceneT<double> c;
  c.mass = 1.0;
  c.spin = s.spin_d;
  c.spin2 = s.spin_d * s.spin_d;
  c.charge2 = s.charge2_d;
  c.chart = s.chart;
  c.r_capture = s.r_capture_d;
  c.r_defer = s.r_horizon_d > 0.0 ? 1.05 * s.r_horizon_d : double(kRingDeferRadius);
  c.r_escape = s.r_escape;
  c.disk_in = s.disk_in_d;
  c.disk_out = s.disk_out_d;
  c.band = s.arming_band;
  c.rtol = s.rtol;
  c.atol = s.atol;
  c.h_min = s.h_min;
  c.h_max = s.h_max;
  c.h_max_per_r = s.h_max_per_r;
  c.tie_tol = s.tie_tol;
  c.drift_limit = s.drift_limit;
  c.guard = s.guard;
  c.safety = s.safety;
  c.min_factor = s.min_factor;
  c.max_factor = s.max_factor;
  c.max_accepted = s.max_accepted;
  c.max_rejects = s.max_rejects;
  c.options = s.options;
  c.reproject_interval = s.reproject_interval;
  return c;
}

enum : uint32_t {
  kTermActive = 0,
  kTermHorizon = 1,
  kTermEscape = 2,
  kTermSingularity = 3,
  kTermStepExhaustion = 4,
  kTermFailure = 5,
  kTermUncertain = 6,
  kTermSurface = 7,
  kTermRejectExhaustion = 8,
  kTermCertified = 9,
};
enum : uint32_t { kEvSing = 1u, kEvHor = 2u, kEvSurf = 4u, kEvEsc = 8u };
// Bits of RayT::armed above the event bits, set once at the ray's start by
// the near-critical test (near_critical.h) from its exact initial state:
// kStartCritical / kStartBranch when the f32 rounding of the initial state
// alone puts the margin within the near-critical / the branch factor of the
// error budget; bits 6-18 the radius r_m of the radial potential's minimum
// and bits 19-31 the margin |mu| there, each as a 13-bit float (r_m: 5
// exponent bits for [2^-2, 2^29), 8 mantissa bits, rounded up, 0 = no test;
// |mu|: 6 exponent bits for [2^-60, 4), 7 mantissa bits, rounded toward
// zero, 0 below 2^-60). The roundings err on the flagging side, and carried
// this way the test costs the trace loop no register.
enum : uint32_t { kStartCritical = 1u << 4, kStartBranch = 1u << 5 };
constexpr uint32_t kStartRmShift = 6u, kStartMuShift = 19u, kStartFieldMask = 0x1fffu;

GL_HD uint32_t gl_f2u(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_uint(v);
#else
  uint32_t u;
  __builtin_memcpy(&u, &v, 4);
  return u;
#endif
}
GL_HD float gl_u2f(uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __uint_as_float(u);
#else
  float v;
  __builtin_memcpy(&v, &u, 4);
  return v;
#endif
}

// r_m as a 13-bit float, rounded up (0 for no test or a non-positive r_m).
GL_HD uint32_t encode_start_radius(float r_m) {
  if (!(r_m > 0.0f)) return 0u;
  const float v = r_m < 0.25f ? 0.25f : (r_m > 5.0e8f ? 5.0e8f : r_m);
  const uint32_t b = gl_f2u(v) + ((1u << 15) - 1u);  // round the mantissa up
  return ((((b >> 23) & 255u) - 124u) << 8) | ((b >> 15) & 255u);
}
GL_HD float decode_start_radius(uint32_t armed) {
  const uint32_t c = (armed >> kStartRmShift) & kStartFieldMask;
  if (c == 0u) return 0.0f;
  return gl_u2f((((c >> 8) + 124u) << 23) | ((c & 255u) << 15));
}

// |mu| as a 13-bit float, rounded toward zero (0 below 2^-60).
GL_HD uint32_t encode_start_margin(float mu) {
  if (!(mu >= 8.673617379884035e-19f)) return 0u;  // 2^-60, and NaN
  const float v = mu > 3.99f ? 3.99f : mu;
  const uint32_t b = gl_f2u(v);
  return ((((b >> 23) & 255u) - 66u) << 7) | ((b >> 16) & 127u);
}
GL_HD float decode_start_margin(uint32_t armed) {
  const uint32_t c = (armed >> kStartMuShift) & kStartFieldMask;
  if (c == 0u) return 0.0f;
  return gl_u2f((((c >> 7) + 66u) << 23) | ((c & 127u) << 16));
}
enum : uint32_t {
  kFlagNonFinite = 1u,
  kFlagRadicand = 2u,
  kFlagDenominator = 4u,
  kFlagDrift = 8u,
  kFlagTurningOrder = 16u,
  kFlagTied = 32u,
  kFlagCertified = 64u,  // horizon proven by the interval capture certificate
  kFlagReconstructed = 128u,  // escape direction interpolated from the stencil
};
enum : uint32_t {
  kOptCertificate = 1u,
  kOptF64Horizon = 2u,
  kOptF64DiskRoot = 4u,
  kOptEscapeMap = 8u,
  kOptRefine = 16u,
  // set by a kernel whose caller asked for the turning-radius plane: every
  // inward-to-outward turning is then localized on the dense output
  kOptTurningRadius = 1u << 30,
};

// Per-ray state. Position and coordinate time / affine parameter carry a
// Kahan compensation term c with true value ~ value - c.
template <class T>
struct RayT {
  T x, y, z, cx, cy, cz;
  T px, py, pz, E;
  T k[7];     // FSAL derivative at the current state: t', x', p'
  T r, rdot;  // radius and dr/dlambda at the current state
  T h;        // signed next step
  T t, ct;    // signed coordinate-time delta
  T lam, clam;
  T lz0, q0;
  T d_null, d_lz, d_q;
  // Smallest inward-to-outward turning radius, -1 if none: localized on the
  // dense output with kOptTurningRadius, else the smaller endpoint radius of
  // the turning step (enough for the near-critical test's proximity check).
  T r_turn;
  // The near-critical test (near_critical.h): d_crit, the largest relative
  // drift sampled around r_m (max(7/8 r_m, (r_m + r_+) / 2) <= r <= 2 r_m),
  // where the capture/scatter decision is taken; r_m and the margin ride in
  // the high bits of `armed`.
  T d_crit;
  uint32_t accepted, rejected, consec;
  uint32_t turnings, crossings;
  int32_t winding;
  uint32_t armed;  // bit per event kind (kEv*)
  uint32_t term, candidates, flags;
  uint32_t polar_side;
  float dir;  // +1 forward, -1 backward affine traversal
};

template <class T>
GL_HD bool finite_val(T v) {
  return v == v && gl_abs<T>(v) <= T(3.0e38);
}

template <class T>
GL_HD T kahan_add(T& sum, T& comp, T v) {
  T yv = v - comp;
  T tv = sum + yv;
  comp = (tv - sum) - yv;
  sum = tv;
  return tv;
}

template <class T>
GL_HD T gl_atan2(T y, T x);
template <>
GL_HD float gl_atan2<float>(float y, float x) {
  return atan2f(y, x);
}
template <>
GL_HD double gl_atan2<double>(double y, double x) {
  return atan2(y, x);
}

template <class T>
GL_HD T pow_m02(T v);  // v^-0.2 for the step controller
template <>
GL_HD float pow_m02<float>(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  // v_log_f32 / v_exp_f32 without the library's rescaling of denormal
  // arguments: the same values for every normal v, and for a denormal v the
  // factor (which the controller clamps to max_factor either way) is +inf
  // instead of a large finite number.
  return __builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf(v));
#else
  return exp2f(-0.2f * log2f(v));
#endif
}
template <>
GL_HD double pow_m02<double>(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  // the step-size factor needs no f64 accuracy
  return double(exp2f(-0.2f * log2f(float(v))));
#else
  return pow(v, -0.2);
#endif
}

template <class T>
GL_HD T gl_log2(T v);
template <>
GL_HD float gl_log2<float>(float v) {
  return log2f(v);
}
template <>
GL_HD double gl_log2<double>(double v) {
  return log2(v);
}
template <class T>
GL_HD T gl_exp2(T v);
template <>
GL_HD float gl_exp2<float>(float v) {
  return exp2f(v);
}
template <>
GL_HD double gl_exp2<double>(double v) {
  return exp2(v);
}

template <class T>
GL_HD T gl_floor(T v);
template <>
GL_HD float gl_floor<float>(float v) {
  return floorf(v);
}
template <>
GL_HD double gl_floor<double>(double v) {
  return floor(v);
}

template <class T>
GL_HD T gl_clamp(T v, T lo, T hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}
// One v_max / v_min (v_med3 for the clamp) instead of a compare and select;
// these return the other operand for a NaN, so the step checks finiteness
// of its new state explicitly.
template <>
GL_HD float gl_clamp<float>(float v, float lo, float hi) {
  return __builtin_fminf(__builtin_fmaxf(v, lo), hi);
}
template <class T>
GL_HD T gl_max(T a, T b) {
  return a > b ? a : b;
}
template <>
GL_HD float gl_max<float>(float a, float b) {
  return __builtin_fmaxf(a, b);
}
template <>
GL_HD double gl_max<double>(double a, double b) {
  return __builtin_fmax(a, b);
}
template <class T>
GL_HD T gl_min(T a, T b) {
  return a < b ? a : b;
}
template <>
GL_HD float gl_min<float>(float a, float b) {
  return __builtin_fminf(a, b);
}
template <>
GL_HD double gl_min<double>(double a, double b) {
  return __builtin_fmin(a, b);
}

// RHS and geometry at (x, p). Returns false on a non-finite or singular
// evaluation.
template <class T>
GL_HD bool eval(const SceneT<T>& sc, T x, T y, T z, T px, T py, T pz, T E, Deriv<T>& d,
                Geom<T>& g) {
  g = geometry<T>(x, y, z, sc.mass, sc.spin, sc.spin2, sc.charge2, sc.chart);
  d = rhs_from_geometry<T>(g, x, y, z, px, py, pz, E, sc.spin, sc.chart);
  return g.ok && finite_val(d.dt) && finite_val(d.dx) && finite_val(d.dy) && finite_val(d.dz) &&
         finite_val(d.dpx) && finite_val(d.dpy) && finite_val(d.dpz);
}

template <class T>
GL_HD T singularity_measure(T r, T spin2, T z) {
  T r2 = r * r;
  return r2 * r2 + spin2 * z * z;
}

// Re-projects the spatial momentum onto the null cone H = 0 at fixed E with
// one Newton step along dH/dp = x', evaluated in double:
//   2H = -E^2 + |p|^2 - f q^2,  p <- p - (2H / (2 |x'|^2)) x'.
template <class T>
GL_HD void reproject_null(const SceneT<T>& sc, const RayT<T>& R, T& px, T& py, T& pz) {
  double x = double(R.x) - double(R.cx), y = double(R.y) - double(R.cy),
         z = double(R.z) - double(R.cz);
  double ppx = double(px), ppy = double(py), ppz = double(pz), E = double(R.E);
  Geom<double> g = geometry<double>(x, y, z, double(sc.mass), double(sc.spin),
                                    double(sc.spin) * double(sc.spin), double(sc.charge2),
                                    double(sc.chart));
  if (!g.ok) return;
  double q = E + g.lx * ppx + g.ly * ppy + g.lz * ppz;
  double fq = g.f * q;
  double vx = ppx - fq * g.lx, vy = ppy - fq * g.ly, vz = ppz - fq * g.lz;
  double h2 = -E * E + ppx * ppx + ppy * ppy + ppz * ppz - g.f * q * q;
  double v2 = vx * vx + vy * vy + vz * vz;
  if (!(v2 > 0.0)) return;
  double a = h2 / (2.0 * v2);
  px = T(ppx - a * vx);
  py = T(ppy - a * vy);
  pz = T(ppz - a * vz);
}

// Evaluates the FSAL derivative at the ray's current state.
template <class T>
GL_HD bool refresh_derivative(const SceneT<T>& sc, RayT<T>& R) {
  Deriv<T> d;
  Geom<T> g;
  if (!eval<T>(sc, R.x, R.y, R.z, R.px, R.py, R.pz, R.E, d, g)) return false;
  R.k[0] = d.dt;
  R.k[1] = d.dx;
  R.k[2] = d.dy;
  R.k[3] = d.dz;
  R.k[4] = d.dpx;
  R.k[5] = d.dpy;
  R.k[6] = d.dpz;
  R.r = g.r;
  R.rdot = radial_rate<T>(g, d);
  return true;
}

template <class T>
GL_HD uint32_t arm_events(const SceneT<T>& sc, T r, T z, uint32_t armed) {
  if (sc.r_capture > T(0) && r - sc.r_capture > sc.band) armed |= kEvHor;
  if (r - sc.r_escape < -sc.band) armed |= kEvEsc;
  if (sc.disk_in > T(0) && gl_abs<T>(z) > sc.band) armed |= kEvSurf;
  // once armed, an event stays armed: the singularity measure is formed only
  // until then
  if (!(armed & kEvSing) && singularity_measure<T>(r, sc.spin2, z) - sc.guard > sc.band)
    armed |= kEvSing;
  return armed;
}

// Initializes a ray from position and covariant spatial momentum (E given).
template <class T>
GL_HD void init_ray(const SceneT<T>& sc, RayT<T>& R, T x, T y, T z, T cx, T cy, T cz, T px,
                    T py, T pz, T E, float dir, uint32_t polar_side) {
  R.x = x;
  R.y = y;
  R.z = z;
  R.cx = cx;
  R.cy = cy;
  R.cz = cz;
  R.px = px;
  R.py = py;
  R.pz = pz;
  R.E = E;
  R.dir = dir;
  R.h = T(dir) * sc.h_min;  // replaced by the caller's initial step
  R.t = R.ct = R.lam = R.clam = T(0);
  R.d_null = R.d_lz = R.d_q = T(0);
  R.accepted = R.rejected = R.consec = 0;
  R.turnings = R.crossings = 0;
  R.winding = 0;
  R.term = kTermActive;
  R.candidates = 0;
  R.flags = 0;
  R.polar_side = polar_side;
  R.r_turn = T(-1);
  R.d_crit = T(0);
  if (!refresh_derivative<T>(sc, R)) {
    R.term = kTermFailure;
    R.flags |= kFlagNonFinite;
    return;
  }
  Geom<T> g = geometry<T>(x, y, z, sc.mass, sc.spin, sc.spin2, sc.charge2, sc.chart);
  R.lz0 = x * py - y * px;
  R.q0 = carter<T>(g, x, y, z, px, py, pz, E, sc.spin2);
  Deriv<T> d;
  d.dt = R.k[0];
  d.dx = R.k[1];
  d.dy = R.k[2];
  d.dz = R.k[3];
  R.d_null = null_residual<T>(d, px, py, pz, E);
  R.armed = arm_events<T>(sc, R.r, z, 0u);
  R.d_crit = R.d_null;
}

// Dense position at fraction th of the step from (x0, y0, z0).
template <class T>
GL_HD void dense_position(const T K[7][7], T h, T x0, T y0, T z0, T th, T& x, T& y, T& z) {
  T w[7];
  dense_weights<T>(th, w);
  T sx = w[0] * K[0][1], sy = w[0] * K[0][2], sz = w[0] * K[0][3];
#pragma unroll
  for (int i = 2; i < 7; ++i) {
    sx += w[i] * K[i][1];
    sy += w[i] * K[i][2];
    sz += w[i] * K[i][3];
  }
  x = x0 + h * sx;
  y = y0 + h * sy;
  z = z0 + h * sz;
}

template <class T>
GL_HD void dense_state(const T K[7][7], T h, const T y0[6], T th, T out[6], T& dtime) {
  T w[7];
  dense_weights<T>(th, w);
  T s[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) s[c] = w[0] * K[0][c];
#pragma unroll
  for (int i = 2; i < 7; ++i) {
#pragma unroll
    for (int c = 0; c < 7; ++c) s[c] += w[i] * K[i][c];
  }
  dtime = h * s[0];
#pragma unroll
  for (int c = 0; c < 6; ++c) out[c] = y0[c] + h * s[c + 1];
}

// Fraction in [0, 1] where r(theta) = target, bracketed by the endpoint
// values g0 = r0 - target and g1 = r1 - target of opposite sign. Illinois
// false position on the dense output.
template <class T, class SC>
GL_HD T radius_root(const SC& sc, const T K[7][7], T h, T x0, T y0, T z0, T g0, T g1, T target) {
  T lo = T(0), hi = T(1), glo = g0, ghi = g1;
  int side = 0;
  T th = T(0.5);
  for (int it = 0; it < (sizeof(T) >= 8 ? 48 : 24); ++it) {
    T den = ghi - glo;
    th = den != T(0) ? lo - glo * (hi - lo) * gl_rcp<T>(den) : T(0.5) * (lo + hi);
    if (!(th > lo && th < hi)) th = T(0.5) * (lo + hi);
    T x, y, z;
    dense_position<T>(K, h, x0, y0, z0, th, x, y, z);
    Geom<T> g = geometry<T>(x, y, z, sc.mass, sc.spin, sc.spin2, sc.charge2, sc.chart);
    T gv = g.r - target;
    if (gv == T(0)) return th;
    if ((gv < T(0)) == (glo < T(0))) {
      lo = th;
      glo = gv;
      if (side == -1) ghi *= T(0.5);
      side = -1;
    } else {
      hi = th;
      ghi = gv;
      if (side == 1) glo *= T(0.5);
      side = 1;
    }
    // bracket in affine parameter: 1e-6 M (per unit of radius beyond 1 M)
    // in f32, 1e-12 M in f64
    if ((hi - lo) * gl_abs<T>(h) <= T(sizeof(T) >= 8 ? 1e-12 : 1e-6) * gl_max<T>(T(1), target)) break;
  }
  return gl_abs<T>(glo) <= gl_abs<T>(ghi) ? lo : hi;
}

// Fraction where z(theta) = 0 on the quartic dense output: bracketed Newton
// on the polynomial, evaluated in the scalar type U (double for the f64 disk
// root solve).
template <class U, class T>
GL_HD U z_root(const T K[7][7], T h, T z0, T z1) {
  // z(theta) = z0 + h sum_i w_i(theta) K_i,z ; coefficients of theta^1..4
  const double d[7][4] = {
      {1.0, -8048581381.0 / 2820520608.0, 8663915743.0 / 2820520608.0,
       -12715105075.0 / 11282082432.0},
      {0, 0, 0, 0},
      {0, 131558114200.0 / 32700410799.0, -68118460800.0 / 10900136933.0,
       87487479700.0 / 32700410799.0},
      {0, -1754552775.0 / 470086768.0, 14199869525.0 / 1410260304.0,
       -10690763975.0 / 1880347072.0},
      {0, 127303824393.0 / 49829197408.0, -318862633887.0 / 49829197408.0,
       701980252875.0 / 199316789632.0},
      {0, -282668133.0 / 205662961.0, 2019193451.0 / 616988883.0, -1453857185.0 / 822651844.0},
      {0, 40617522.0 / 29380423.0, -110615467.0 / 29380423.0, 69997945.0 / 29380423.0}};
  U c[4] = {U(0), U(0), U(0), U(0)};
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    if (i == 1) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) c[j] += U(d[i][j]) * U(K[i][3]);
  }
  U hh = U(h), zz0 = U(z0);
  for (int j = 0; j < 4; ++j) c[j] *= hh;
  // the derivative's coefficients, and the quotients: the f32 trace divides
  // through the hardware reciprocal (1 ulp) as everywhere else in its step,
  // the f64 disk root keeps the correctly rounded division
  const U dc[3] = {U(2) * c[1], U(3) * c[2], U(4) * c[3]};
  auto quot = [](U a, U b) -> U {
    if constexpr (sizeof(U) >= sizeof(double)) {
      return a / b;
    } else {
      return a * gl_rcp<U>(b);
    }
  };
  U lo = U(0), hi = U(1), flo = zz0;
  U th = quot(flo, flo - U(z1));
  if (!(th > lo && th < hi)) th = U(0.5);
  // Bracket width that ends the iteration: 1e-13 in f64; in f32 two ulps of
  // a fraction near 1 (2^-22). A 1e-13 bracket lies below binary32
  // resolution, so in f32 it ended only at the iteration cap whenever the
  // polynomial never evaluated to exactly zero (about one root in ten),
  // with no gain in accuracy: the root is already within 2e-7 of the f64
  // root when the bracket reaches 2^-22.
  const U stop = sizeof(U) >= sizeof(double) ? U(1e-13) : U(2.384185791015625e-7);
  for (int it = 0; it < 40; ++it) {
    U f = zz0 + th * (c[0] + th * (c[1] + th * (c[2] + th * c[3])));
    U df = c[0] + th * (dc[0] + th * (dc[1] + th * dc[2]));
    if (f == U(0)) return th;
    if ((f < U(0)) == (flo < U(0))) {
      lo = th;
      flo = f;
    } else {
      hi = th;
    }
    U nt = th - quot(f, df);
    th = (nt > lo && nt < hi) ? nt : U(0.5) * (lo + hi);
    if (hi - lo <= stop) break;
  }
  return th;
}

// The radial turning inside an accepted step: the root of the radial
// traversal velocity dir * grad r . dx/dlambda on the dense output, with
// dx/dlambda(theta) = sum_i w_i'(theta) K_i (geodesic.rs:330-359 localizes
// the same sign change of the radial velocity). The endpoint velocities
// v0, v1 (traversal sign applied) bracket it. Illinois false position to a
// bracket of a quarter of the tie tolerance in affine parameter; returns the
// bracket [lo, hi] and the radius at the last point evaluated inside it.
template <class T, class SC>
GL_HD void turning_root(const SC& sc, const T K[7][7], T h, T x0, T y0, T z0, T v0, T v1, T dir,
                        T& lo_out, T& hi_out, T& r_out) {
  T lo = T(0), hi = T(1), glo = v0, ghi = v1;
  int side = 0;
  T r_at = T(0);
  const T stop = T(0.25) * sc.tie_tol;
  const T hmag = gl_abs<T>(h);
  for (int it = 0; it < 32; ++it) {
    const T den = ghi - glo;
    T th = den != T(0) ? lo - glo * (hi - lo) * gl_rcp<T>(den) : T(0.5) * (lo + hi);
    if (!(th > lo && th < hi)) th = T(0.5) * (lo + hi);
    T x, y, z;
    dense_position<T>(K, h, x0, y0, z0, th, x, y, z);
    T w[7];
    dense_slopes<T>(th, w);
    T vx = w[0] * K[0][1], vy = w[0] * K[0][2], vz = w[0] * K[0][3];
#pragma unroll
    for (int i = 2; i < 7; ++i) {
      vx += w[i] * K[i][1];
      vy += w[i] * K[i][2];
      vz += w[i] * K[i][3];
    }
    const Geom<T> g = geometry<T>(x, y, z, sc.mass, sc.spin, sc.spin2, sc.charge2, sc.chart);
    const T gv = dir * (g.grx * vx + g.gry * vy + g.grz * vz);
    r_at = g.r;
    if (gv == T(0)) {
      lo = hi = th;
      break;
    }
    if ((gv < T(0)) == (glo < T(0))) {
      lo = th;
      glo = gv;
      if (side == -1) ghi *= T(0.5);
      side = -1;
    } else {
      hi = th;
      ghi = gv;
      if (side == 1) glo *= T(0.5);
      side = 1;
    }
    if ((hi - lo) * hmag <= stop) break;
  }
  lo_out = lo;
  hi_out = hi;
  r_out = r_at;
}

template <class T>
GL_HD void note_turning_radius(RayT<T>& R, T r) {
  R.r_turn = R.r_turn < T(0) ? r : gl_min<T>(R.r_turn, r);
}

template <class T>
GL_HD T step_cap(const SceneT<T>& sc, T r) {
  T cap = sc.h_max;
  if (sc.h_max_per_r > T(0)) cap = gl_min<T>(cap, sc.h_max_per_r * r);
  return gl_max<T>(cap, sc.h_min);
}

// Commits the relative drifts of the invariants at a state whose RHS d and
// geometry g are known (geodesic.rs:575-598; Carter in the Cartesian form).
template <class T>
GL_HD void update_drift(const SceneT<T>& sc, RayT<T>& R, const Geom<T>& g, const Deriv<T>& d,
                        T x, T y, T z, T px, T py, T pz) {
  T nr = null_residual<T>(d, px, py, pz, R.E);
  T lz = x * py - y * px;
  T q = carter<T>(g, x, y, z, px, py, pz, R.E, sc.spin2);
  T dl = gl_abs<T>(lz - R.lz0) * gl_rcp<T>(gl_max<T>(T(1), gl_abs<T>(R.lz0)));
  T dq = gl_abs<T>(q - R.q0) * gl_rcp<T>(gl_max<T>(T(1), gl_abs<T>(R.q0)));
  R.d_null = gl_max<T>(R.d_null, nr);
  R.d_lz = gl_max<T>(R.d_lz, dl);
  R.d_q = gl_max<T>(R.d_q, dq);
  // the near-critical budget counts the drift where the capture/scatter
  // decision is taken, around r_m: up to 2 r_m (beyond, the invariants' own
  // evaluation, terms of size r^2 p^2 cancelling to Carter's constant, is
  // the larger error) and down to 7/8 r_m, for a ray near the critical curve
  // orbits at r ~ r_m on either side of it (its constants differ from the
  // exact ray's by the drift, and the carried r_m is rounded up by up to
  // 2^-8; a band starting at r_m missed the orbits of rays that circled ~30
  // times and left with the wrong class), but not below halfway to the
  // horizon: there the ray is decided, and in a chart singular on the
  // horizon its momentum, and with it the evaluation error of Carter's
  // constant, grows without bound.
  const T r_m = T(decode_start_radius(R.armed));
  const T r_lo = gl_max<T>(T(0.875) * r_m, T(0.5) * (r_m + sc.r_capture));
  if (g.r >= r_lo && g.r <= T(2) * r_m)
    R.d_crit = gl_max<T>(R.d_crit, gl_max<T>(nr, gl_max<T>(dl, dq)));
  if (!(R.d_null <= sc.drift_limit && R.d_lz <= sc.drift_limit && R.d_q <= sc.drift_limit))
    R.flags |= kFlagDrift;
}

// Signed azimuth winding (geodesic.rs:406-418): the reference counts a jump
// of atan2(y, x) by more than pi between committed states, i.e. a crossing of
// the branch cut, the half-plane y = 0, x < 0. Here the crossing is read off
// the chord between the two committed states without atan2: y changes sign
// and the chord meets y = 0 at x < 0. Going from y >= 0 to y < 0 there
// (atan2 jumps from +pi to -pi) adds one turn, the reverse removes one.
// Chords span at most a fraction of a turn (|h| <= h_max_per_r r), where
// both tests agree.
template <class T>
GL_HD int32_t winding_step(T x0, T y0, T x1, T y1) {
  const bool up0 = y0 >= T(0), up1 = y1 >= T(0);
  // x at the crossing, scaled by (y0 - y1) to avoid a division
  const T xc = x0 * (y0 - y1) + (x1 - x0) * y0;
  const bool neg = (y0 - y1) > T(0) ? xc < T(0) : xc > T(0);
  return (up0 != up1 && neg) ? (up0 ? 1 : -1) : 0;
}

// Source anchor (r, phi_s) and frequency ratio g of the prograde Keplerian
// emitter hit at (x, y, r) (surface.rs:75-146, docs/physics.md:512-556).
// Returns false when the circular orbit does not exist or is not timelike.
template <class U>
GL_HD bool surface_observable(U mass, U spin, U charge2, U chart, U omega, U x, U y, U r, U E,
                              U lz, U out[3]) {
  U crs = mass * r - charge2;
  if (!(crs >= U(0))) return false;
  U cr = gl_sqrt<U>(crs);
  U bs = spin < U(0) ? U(-1) : U(1);
  U r2 = r * r;
  U den = r2 + bs * spin * cr;
  if (!(den > U(0))) return false;
  U om = bs * cr / den;
  U rn = U(2) * mass * r - charge2;
  U a2 = spin * spin;
  U tdisc = r2 - U(3) * mass * r + U(2) * charge2 + U(2) * bs * spin * cr;
  if (!(tdisc > U(0))) return false;
  U delta = r2 - U(2) * mass * r + a2 + charge2;
  U gtt = U(-1) + rn / r2;
  U gtp = -rn * spin / r2;
  U rf = r2 + a2;
  U gpp = (rf * rf - a2 * delta) / r2;
  U cn = gtt + U(2) * gtp * om + gpp * om * om;
  if (!(cn < U(0))) return false;
  U ut = U(1) / gl_sqrt<U>(-cn);
  U ef = ut * (E - om * lz);
  if (!(ef > U(0))) return false;
  U g = omega / ef;
  U raw = gl_atan2<U>(y, x) - gl_atan2<U>(chart * spin, r);
  const U pi = U(3.14159265358979323846), tau = U(6.28318530717958647692);
  U phi = raw - tau * gl_floor<U>((raw + pi) / tau);
  out[0] = r;
  out[1] = phi;
  out[2] = g;
  return true;
}

// One adaptive DP5 attempt. On acceptance the ray advances, events are
// detected and localized, and the branch key, drifts and arming are
// committed; on termination R.term becomes nonzero and (x, p) hold the
// localized terminal state.
//
// The attempt is straight-line code up to its outcome, and the ray state is
// committed with selects: an early return (step exhaustion, a non-finite
// stage, a rejection) would make the compiler merge the old and the new
// state at the join, and it does so through a third set of registers, with
// a copy of every state value on both paths. Only rare work sits behind
// branches: the event roots, the terminal localization, the drift sample and
// the re-projection. Step exhaustion is decided before the attempt, whose
// arithmetic is then discarded.
template <class T>
GL_HD void step(const SceneT<T>& sc, RayT<T>& R, T omega, bool reproject = false,
                 bool drift_now = true) {
  const bool exhausted = R.accepted >= sc.max_accepted;
  const T h = R.h;
  T K[7][7];
#pragma unroll
  for (int c = 0; c < 7; ++c) K[0][c] = R.k[c];
  const T y0[6] = {R.x, R.y, R.z, R.px, R.py, R.pz};
  const T A[6][6] = {
      {T(GL_A21), T(0), T(0), T(0), T(0), T(0)},
      {T(GL_A31), T(GL_A32), T(0), T(0), T(0), T(0)},
      {T(GL_A41), T(GL_A42), T(GL_A43), T(0), T(0), T(0)},
      {T(GL_A51), T(GL_A52), T(GL_A53), T(GL_A54), T(0), T(0)},
      {T(GL_A61), T(GL_A62), T(GL_A63), T(GL_A64), T(GL_A65), T(0)},
      {T(GL_B1), T(0), T(GL_B3), T(GL_B4), T(GL_B5), T(GL_B6)}};
  bool ok = true;
  Deriv<T> d;
  Geom<T> g;
  T xn[3], cn[3], pn[3];
#pragma unroll
  for (int s = 1; s < 7; ++s) {
    T ys[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      T acc = A[s - 1][0] * K[0][c + 1];
#pragma unroll
      for (int j = 1; j < s; ++j)
        if (!(s == 6 && j == 1)) acc += A[s - 1][j] * K[j][c + 1];
      ys[c] = y0[c] + h * acc;
      if (s == 6) {
        if (c < 3) {
          // compensated position update
          T cc = c == 0 ? R.cx : (c == 1 ? R.cy : R.cz);
          T yk = h * acc - cc;
          T tn = y0[c] + yk;
          cn[c] = (tn - y0[c]) - yk;
          xn[c] = tn;
          ys[c] = tn;
        } else {
          pn[c - 3] = ys[c];
        }
      }
    }
    // Intermediate stages need only the derivative; the endpoint (s = 6)
    // also the geometry. Finiteness of every stage is checked once below,
    // through the error estimate and the new state.
    if (s < 6) {
      ok = rhs_only<T>(ys[0], ys[1], ys[2], ys[3], ys[4], ys[5], R.E, sc.mass, sc.spin, sc.spin2,
                       sc.charge2, sc.chart, d) && ok;
    } else {
      g = geometry<T>(ys[0], ys[1], ys[2], sc.mass, sc.spin, sc.spin2, sc.charge2, sc.chart);
      d = rhs_from_geometry<T>(g, ys[0], ys[1], ys[2], ys[3], ys[4], ys[5], R.E, sc.spin,
                               sc.chart);
      ok = g.ok && ok;
    }
    K[s][0] = d.dt;
    K[s][1] = d.dx;
    K[s][2] = d.dy;
    K[s][3] = d.dz;
    K[s][4] = d.dpx;
    K[s][5] = d.dpy;
    K[s][6] = d.dpz;
  }
  const T Ew[7] = {T(GL_E1), T(0), T(GL_E3), T(GL_E4), T(GL_E5), T(GL_E6), T(GL_E7)};
  T err[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) {
    T acc = Ew[0] * K[0][c];
#pragma unroll
    for (int j = 2; j < 7; ++j) acc += Ew[j] * K[j][c];
    err[c] = h * acc;
  }
  T dtinc = h * (T(GL_B1) * K[0][0] + T(GL_B3) * K[2][0] + T(GL_B4) * K[3][0] +
                 T(GL_B5) * K[4][0] + T(GL_B6) * K[5][0]);
  // Group error norm (policy.rs:137-158) with reciprocal scales.
  T en = gl_abs<T>(err[0]) *
         gl_rcp<T>(sc.atol + sc.rtol * gl_max<T>(gl_abs<T>(R.t), gl_abs<T>(R.t + dtinc)));
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    T e1 = gl_abs<T>(err[c + 1]) *
           gl_rcp<T>(sc.atol + sc.rtol * gl_max<T>(gl_abs<T>(y0[c]), gl_abs<T>(xn[c])));
    T e2 = gl_abs<T>(err[c + 4]) *
           gl_rcp<T>(sc.atol + sc.rtol * gl_max<T>(gl_abs<T>(y0[c + 3]), gl_abs<T>(pn[c])));
    en = gl_max<T>(en, gl_max<T>(e1, e2));
  }
  // every stage, the error estimate and the new state must be finite (the
  // max/min above drop NaN operands; a non-finite stage reaches the error or
  // the new state through its weights or through the later stages)
  const T probe = xn[0] + xn[1] + xn[2] + pn[0] + pn[1] + pn[2] + dtinc + err[0] + err[1] +
                  err[2] + err[3] + err[4] + err[5] + err[6];
  const bool fail = !ok || !(en == en) || en > T(3.0e38) || !finite_val<T>(probe);
  const bool live = !exhausted && !fail;
  const bool reject = live && en > T(1);
  const bool accept = live && !(en > T(1));
  const T hmag = gl_abs<T>(h);
  const T dir = T(R.dir);
  // The controller factor 0.9 E^-1/5 (policy.rs:160-173), shared by both
  // outcomes.
  const T pw = sc.safety * pow_m02<T>(en);

  // Rejection (policy.rs:137-173): the reject limit, then the minimum step.
  const uint32_t consec_rej = R.consec + 1u;
  const bool rej_exhausted = reject && consec_rej >= sc.max_rejects;
  const bool rej_failed = reject && !rej_exhausted && hmag <= sc.h_min;
  const T f_rej = gl_clamp<T>(pw, sc.min_factor, T(1));
  const T h_rej = dir * gl_clamp<T>(hmag * f_rej, sc.h_min, step_cap<T>(sc, R.r));

  const bool after_reject = R.consec != 0;
  const T r0 = R.r, r1 = g.r;
  const T rd1 = radial_rate<T>(g, d);
  const T z0 = y0[2], z1 = xn[2];
  const T v0 = dir * R.rdot, v1 = dir * rd1;
  const bool turning = (v0 < T(0) && v1 >= T(0)) || (v0 > T(0) && v1 <= T(0));
  const bool zcross = (z0 < T(0) && z1 >= T(0)) || (z0 > T(0) && z1 <= T(0));

  // Terminal event candidates of an accepted attempt (fraction > 1 means
  // none).
  T fr[4] = {T(2), T(2), T(2), T(2)};
  // the measure r^4 + a^2 z^2 falls to the guard only where r^4 <= guard
  if (accept && (R.armed & kEvSing) && (r1 * r1) * (r1 * r1) <= sc.guard) {
    T m0 = singularity_measure<T>(r0, sc.spin2, z0) - sc.guard;
    T m1 = singularity_measure<T>(r1, sc.spin2, z1) - sc.guard;
    if (m0 > T(0) && m1 <= T(0)) fr[0] = gl_clamp<T>(m0 * gl_rcp<T>(m0 - m1), T(0), T(1));
  }
  if (accept && (R.armed & kEvHor) && r0 > sc.r_capture && r1 <= sc.r_capture)
    fr[1] = radius_root<T>(sc, K, h, y0[0], y0[1], y0[2], r0 - sc.r_capture, r1 - sc.r_capture,
                           sc.r_capture);
  T thz = T(2);
  if (accept && zcross && (R.armed & kEvSurf)) {
    thz = z_root<T, T>(K, h, z0, z1);
    T xs, ysv, zs;
    dense_position<T>(K, h, y0[0], y0[1], y0[2], thz, xs, ysv, zs);
    // On the equatorial plane the oblate radius has the closed form
    // r^2 = x^2 + y^2 - a^2 (the quartic's root at z = 0); the disk test
    // compares squares, and a crossing inside the ring (r^2 < 0) fails it.
    const T r2s = xs * xs + ysv * ysv - sc.spin2;
    if (r2s >= sc.disk_in * sc.disk_in && r2s <= sc.disk_out * sc.disk_out) fr[2] = thz;
  }
  if (accept && (R.armed & kEvEsc) && r0 < sc.r_escape && r1 >= sc.r_escape)
    fr[3] = radius_root<T>(sc, K, h, y0[0], y0[1], y0[2], r0 - sc.r_escape, r1 - sc.r_escape,
                           sc.r_escape);
  int kind = 0;
  T best = fr[0];
#pragma unroll
  for (int i = 1; i < 4; ++i)
    if (fr[i] < best) {
      best = fr[i];
      kind = i;
    }

  const bool terminal = best <= T(1);  // only an accepted attempt has candidates
  const bool commit = accept && !terminal;

  // Counters and outcome of the attempt.
  R.accepted += accept ? 1u : 0u;
  R.rejected += reject ? 1u : 0u;
  R.consec = reject ? consec_rej : (accept ? 0u : R.consec);
  uint32_t term = R.term;
  if (exhausted) term = kTermStepExhaustion;
  if (!exhausted && fail) term = kTermFailure;
  if (rej_exhausted) term = kTermRejectExhaustion;
  if (rej_failed) term = kTermFailure;
  R.term = term;
  // A non-finite stage is a typed numerical failure (integrator.rs:170-173).
  R.flags |= (!exhausted && fail) ? kFlagNonFinite : 0u;

  // What a non-terminal accepted attempt reads of the new state, before the
  // state is committed: the drift sample, the radial turning, the crossing
  // and the winding.
  if (commit && (reproject || drift_now))
    update_drift<T>(sc, R, g, d, xn[0], xn[1], xn[2], pn[0], pn[1], pn[2]);
  if (commit && turning && v0 < T(0)) {
    T rt = gl_min<T>(r0, r1);
    if (sc.options & kOptTurningRadius) {
      T tlo, thi;
      turning_root<T>(sc, K, h, y0[0], y0[1], y0[2], v0, v1, dir, tlo, thi, rt);
    }
    note_turning_radius<T>(R, rt);
  }
  R.turnings += (commit && turning) ? 1u : 0u;
  R.crossings += (commit && zcross) ? 1u : 0u;
  int32_t wind = commit ? winding_step<T>(y0[0], y0[1], xn[0], xn[1]) : 0;

  // The state the attempt moves to: the step's end, or the localized event.
  T ox = xn[0], oy = xn[1], oz = xn[2];
  T ocx = cn[0], ocy = cn[1], ocz = cn[2];
  T opx = pn[0], opy = pn[1], opz = pn[2];
  T orr = r1, add_t = dtinc, add_lam = h;
  if (terminal) {
    // An accepted attempt that ends the ray on its earliest event.
    uint32_t mask = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (fr[i] <= T(1) && gl_abs<T>(fr[i] - best) * hmag <= sc.tie_tol) mask |= 1u << i;
    if (__builtin_popcount(mask) > 1) R.flags |= kFlagTied;
    if (zcross && thz > T(1)) thz = z_root<T, T>(K, h, z0, z1);
    T ys[6], dtl;
    dense_state<T>(K, h, y0, best, ys, dtl);
    T cl[3] = {T(0), T(0), T(0)};
    if (kind == 2 && (sc.options & kOptF64DiskRoot)) {
      // f64 root of z(theta) on the quartic dense output and the localized
      // position in f64, kept as value + compensation.
      double thd = z_root<double, T>(K, h, z0, z1);
      double w[7];
      dense_weights<double>(thd, w);
      const double cc[3] = {double(R.cx), double(R.cy), double(R.cz)};
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        double acc = w[0] * double(K[0][c + 1]);
#pragma unroll
        for (int i = 2; i < 7; ++i) acc += w[i] * double(K[i][c + 1]);
        double v = double(y0[c]) + double(h) * acc - (c < 3 ? cc[c] : 0.0);
        ys[c] = T(v);
        if (c < 3) cl[c] = T(double(ys[c]) - v);
      }
      thz = T(thd);
      best = T(thd);
    }
    add_t = dtl;
    add_lam = best * h;
    Deriv<T> dl;
    Geom<T> gl_;
    bool okl = eval<T>(sc, ys[0], ys[1], ys[2], ys[3], ys[4], ys[5], R.E, dl, gl_);
    if (turning) {
      // The turning commits only when its root bracket lies strictly before
      // the terminal fraction; a terminal inside the bracket leaves the
      // order of the two unresolved (geodesic_integration.wgsl:298-318).
      T tlo, thi, rt;
      turning_root<T>(sc, K, h, y0[0], y0[1], y0[2], v0, v1, dir, tlo, thi, rt);
      if (thi < best) {
        R.turnings += 1;
        if (v0 < T(0)) note_turning_radius<T>(R, rt);
      } else if (tlo < best) {
        R.flags |= kFlagTurningOrder;
      }
    }
    if (zcross && thz < best) R.crossings += 1;
    wind = winding_step<T>(y0[0], y0[1], ys[0], ys[1]);
    if (okl)
      update_drift<T>(sc, R, gl_, dl, ys[0], ys[1], ys[2], ys[3], ys[4], ys[5]);
    else
      R.flags |= kFlagNonFinite;
    const uint32_t terms[4] = {kTermSingularity, kTermHorizon, kTermSurface, kTermEscape};
    R.term = terms[kind];
    R.candidates = mask;
    ox = ys[0];
    oy = ys[1];
    oz = ys[2];
    ocx = cl[0];
    ocy = cl[1];
    ocz = cl[2];
    opx = ys[3];
    opy = ys[4];
    opz = ys[5];
    orr = gl_.r;
  }

  // Commit by selects: an accepted attempt moves the ray (a rejected or
  // failed one keeps it and takes h_rej).
  R.winding += wind;
  {
    T t = R.t, ct = R.ct, lam = R.lam, clam = R.clam;
    kahan_add<T>(t, ct, add_t);
    kahan_add<T>(lam, clam, add_lam);
    R.t = accept ? t : R.t;
    R.ct = accept ? ct : R.ct;
    R.lam = accept ? lam : R.lam;
    R.clam = accept ? clam : R.clam;
  }
  R.x = accept ? ox : R.x;
  R.y = accept ? oy : R.y;
  R.z = accept ? oz : R.z;
  R.cx = accept ? ocx : R.cx;
  R.cy = accept ? ocy : R.cy;
  R.cz = accept ? ocz : R.cz;
  R.px = accept ? opx : R.px;
  R.py = accept ? opy : R.py;
  R.pz = accept ? opz : R.pz;
  R.r = accept ? orr : R.r;
#pragma unroll
  for (int c = 0; c < 7; ++c) R.k[c] = commit ? K[6][c] : R.k[c];
  R.rdot = commit ? rd1 : R.rdot;
  R.armed = commit ? arm_events<T>(sc, r1, z1, R.armed) : R.armed;
  // Step control of policy.rs:160-173 (safety 0.9, factor in [0.2, 5] by
  // default: the scene's safety, min_factor, max_factor) with
  // no growth on the step right after a rejection (Hairer, Norsett & Wanner,
  // Solving ODEs I, II.4), which breaks the accept/reject alternation of
  // rays falling inward.
  {
    T f = en == T(0) ? sc.max_factor : pw;
    f = gl_clamp<T>(f, sc.min_factor, after_reject ? T(1) : sc.max_factor);
    const T h_acc = dir * gl_clamp<T>(hmag * f, sc.h_min, step_cap<T>(sc, r1));
    R.h = commit ? h_acc : (reject ? h_rej : R.h);
  }
  if (commit && reproject) {
    reproject_null<T>(sc, R, R.px, R.py, R.pz);
    // The radial rate the next step compares against stays the committed
    // one: close to a turning point the radial velocity is tiny, and the
    // re-projected momentum can flip its sign without any step, so a turning
    // would pass between two steps unseen.
    const T rdot_committed = R.rdot;
    if (!refresh_derivative<T>(sc, R)) {
      R.term = kTermFailure;
      R.flags |= kFlagNonFinite;
    }
    R.rdot = rdot_committed;
  }
}

}  // namespace gl
