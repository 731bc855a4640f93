// Device-side pieces shared by the trace kernels and the adjoint kernels:
// the escape-map view, the deferred-ray buffer, pixel mapping, the initial
// ray, and the output writer.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gravlume.h"
#include "certificate.h"
#include "gl_types.h"
#include "near_critical.h"
#include "trace_device.h"

namespace gl {

struct MapView {
  const float* rgb;  // 3 planes of w*h texels
  uint32_t w, h;
};

// Deferred near-horizon rays: full f32 state plus where their outputs go.
struct DeferRecord {
  RayT<float> ray;
  uint32_t frame_scene;
  uint32_t index;
  uint32_t bucket;  // the finisher's claim-order bucket (defer_bucket)
};

// Device workspace of one launch: the deferred near-horizon rays, the
// near-critical rays for the f64 refine pass, their claim orders and their
// counters (kCounterWords words): count[0] deferred records, [1] the
// finisher's claim cursor, [2] refine records, [3] the refine pass's claim
// cursor, [kBatchCursor] the f64 batch's claim cursor,
// [kDeferHist + b] / [kRefineHist + b] records in claim bucket b,
// [kDeferFill + b] / [kRefineFill + b] the ordering passes' cursors.
constexpr uint32_t kBatchCursor = 4u;  // the f64 geodesic batch's claim cursor
constexpr uint32_t kClaimBuckets = 32u;
constexpr uint32_t kDeferHist = 32u;
constexpr uint32_t kDeferFill = kDeferHist + kClaimBuckets;
constexpr uint32_t kRefineHist = kDeferFill + kClaimBuckets;
constexpr uint32_t kRefineFill = kRefineHist + kClaimBuckets;
constexpr uint32_t kCounterWords = kRefineFill + kClaimBuckets + 32u;

struct RefineRecord;
struct DeferBuf {
  DeferRecord* rec;  // nullptr: no f64 near-horizon deferral in this launch
  unsigned int* count;
  uint32_t capacity;
  uint32_t refine_capacity;
  RefineRecord* refine;    // nullptr: no refine pass in this launch
  uint32_t* order;         // the finisher's claim order over rec (capacity entries)
  uint32_t* refine_order;  // the refine pass's claim order (refine_capacity entries)
};

// The f64 passes claim their longest chains first (longest-processing-time
// order), by bucket, bucket 0 first. The finisher groups deferred rays by
// the radius r_m of their radial potential's minimum: the closer r_m lies to
// the horizon, the longer the ray orbits inside the f64 island (on the
// a = 0.998 near-critical batch rays with r_m < 1.09 take ~850 f64 attempts,
// rays with r_m > 2 two). Bucket b covers log2 r_m in [b / 16, (b + 1) / 16);
// r_m < 1 falls into bucket 0, r_m >= 4 and rays without the test (r_m = 0)
// into the last.
__device__ __forceinline__ uint32_t defer_bucket(float r_m) {
  if (!(r_m > 0.0f)) return kClaimBuckets - 1u;
  const float v = 16.0f * __log2f(fmaxf(r_m, 1.0f));
  return v >= float(kClaimBuckets - 1u) ? kClaimBuckets - 1u : uint32_t(v);
}

// The refine pass groups rays by the attempts their f32 trace took (its f64
// trace at rtol 1e-12 takes ~16 times as many, correlated at 0.98 in log on
// the near-critical batch): bucket 31 - floor(2 log2(attempts + 1)), so
// bucket 0 holds rays of 2^15.5 attempts and more.
__device__ __forceinline__ uint32_t refine_bucket(uint32_t attempts) {
  const float v = 2.0f * __log2f(float(attempts) + 1.0f);
  return v >= float(kClaimBuckets - 1u) ? 0u : kClaimBuckets - 1u - uint32_t(v);
}

// Clears the workspace counters; called by every lane of one wavefront.
__device__ __forceinline__ void clear_counters(unsigned int* count) {
  for (uint32_t w = threadIdx.x; w < kCounterWords; w += blockDim.x) count[w] = 0u;
}

// ---------------------------------------------------------------------------
// Device helpers

__device__ __forceinline__ void image_pixel(uint32_t item, uint32_t tiles_x, uint32_t& px,
                                            uint32_t& py) {
  uint32_t tile = item >> 6, t = item & 63u;
  px = (tile % tiles_x) * 8u + (t & 7u);
  py = (tile / tiles_x) * 8u + (t >> 3);
}

// Initial state (t, x, y, z, p_t, p_x, p_y, p_z) in f64 of the backward ray
// through viewport point (px + ox, py + oy): the sight-plane point
// (view.rs:107-123), the direction d = (s_x e_R + s_y e_U - e_A)/|.|
// from the f64 tetrad, p^mu = omega (u - d), lowered with the Kerr-Schild
// metric g = eta + f l l at the observer (scene.rs:211-271).
__device__ __forceinline__ void image_state_f64(const SceneDev& s, uint32_t px, uint32_t py,
                                                double ox, double oy, double y[8]) {
  const double nx = 2.0 * (double(px) + ox) / double(s.width) - 1.0;
  const double ny = 1.0 - 2.0 * (double(py) + oy) / double(s.height);
  const double sx = s.aspect_d * s.tan_half_fov_d * nx;
  const double sy = s.tan_half_fov_d * ny;
  const double inv = gl_rsqrt<double>(1.0 + sx * sx + sy * sy);
  double d[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) d[i] = (sx * s.right_d[i] + sy * s.up_d[i] - s.arrival_d[i]) * inv;
  const double om = double(s.omega);
  const double pu[4] = {om * (s.u_t_d - d[0]), -om * d[1], -om * d[2], -om * d[3]};
  const Geom<double> g = geometry<double>(s.obs_d[0], s.obs_d[1], s.obs_d[2], 1.0, s.spin_d,
                                          s.spin_d * s.spin_d, s.charge2_d, double(s.chart));
  // l_mu p^mu with the covariant principal null vector l_mu = (1, l)
  const double lp = pu[0] + g.lx * pu[1] + g.ly * pu[2] + g.lz * pu[3];
  const double fl = g.f * lp;
  y[0] = 0.0;
  y[1] = s.obs_d[0];
  y[2] = s.obs_d[1];
  y[3] = s.obs_d[2];
  y[4] = -pu[0] + fl;
  y[5] = pu[1] + fl * g.lx;
  y[6] = pu[2] + fl * g.ly;
  y[7] = pu[3] + fl * g.lz;
}

// The start of the near-critical test (near_critical.h) from a ray's exact
// f64 initial state y: the radius r_m of its radial potential's minimum
// (0 when the scene does not test), the margin |mu| there, and the start
// bits (kStartCritical, kStartBranch) when the f32 rounding of y alone
// (f32_start_error; skipped for an exact f64 start) puts |mu| within the
// near-critical or the branch factor. The trace carries these to its end,
// where ray_verdict adds its own drift near r_m.
struct StartTest {
  uint32_t bits;  // kStartCritical | kStartBranch | r_m and |mu| (trace_device.h)
};

__device__ __forceinline__ StartTest start_test(const SceneDev& s, const double* y, bool exact) {
  StartTest t{0u};
  if (!(s.near_critical_factor > 0.0f) || !(s.r_horizon_d > 0.0)) return t;
  const MarginAndBound mb =
      state_margin_and_bound(y, s.spin_d, s.charge2_d, double(s.chart), s.r_horizon_d);
  const RadialMargin<double>& m = mb.m;
  if (!m.ok) return t;
  const double mu = fabs(m.mu);
  t.bits = (encode_start_radius(float(m.r_m)) << kStartRmShift) |
           (encode_start_margin(float(mu)) << kStartMuShift);
  // the start error can flag the ray only within the branch factor of its
  // bound (start_error_bound_from >= f32_start_error); far from the critical
  // curve it is not evaluated
  if (!exact && mu <= double(s.branch_factor) * mb.bound) {
    double e0 = f32_start_error(y, s.spin_d, s.charge2_d, double(s.chart));
    if (!(e0 == e0)) e0 = 1.0;
    if (mu <= double(s.near_critical_factor) * e0) t.bits |= kStartCritical;
    if (mu <= double(s.branch_factor) * e0) t.bits |= kStartBranch;
  }
  return t;
}

// The f32 image ray: the rounding of the f64 state above, and its start
// test. Out of line and returned in registers (five dwords), so the ray
// state of the calling kernel never has its address taken and stays in
// VGPRs.
struct Momentum {
  float px, py, pz, E;
  StartTest start;
};
static_assert(sizeof(Momentum) == 20, "five dwords in registers");

__device__ __noinline__ Momentum image_momentum(const SceneDev& s, uint32_t px, uint32_t py,
                                                float ox, float oy) {
  double y[8];
  image_state_f64(s, px, py, double(ox), double(oy), y);
  return Momentum{float(y[5]), float(y[6]), float(y[7]), float(-y[4]), start_test(s, y, false)};
}

__device__ __noinline__ StartTest batch_start_test(const SceneDev& s, const double* y8) {
  double y[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) y[i] = y8[i];
  return start_test(s, y, false);
}

template <class T>
__device__ __forceinline__ void apply_start(RayT<T>& R, const StartTest& t) {
  R.armed |= t.bits;
}

// The interval radial-capture certificate (certificate.h) of a backward ray
// starting at (x, y, z), from its constants of motion there. Out of line,
// once per ray.
__device__ __noinline__ bool certify_capture(const SceneDev& s, float x, float y, float z,
                                             float px, float py, float pz, float E, float r_start,
                                             float dr_future) {
  const Geom<float> g = geometry<float>(x, y, z, s.mass, s.spin, s.spin2, s.charge2, s.chart);
  const float L = x * py - y * px;
  const float Q = carter<float>(g, x, y, z, px, py, pz, E, s.spin2);
  const float q = float(sqrt(s.charge2_d));
  return capture_certified(s.mass, s.spin, q, s.r_horizon, r_start, x * x + y * y, dr_future, E,
                           L, Q);
}

// The backward ray through point (px + ox, py + oy) of the viewport, with
// the capture certificate applied when the scene enables it.
__device__ __forceinline__ void init_image_ray_at(const SceneDev& s, const SceneT<float>& sc,
                                                  uint32_t px, uint32_t py, float ox, float oy,
                                                  RayT<float>& R) {
  Momentum m = image_momentum(s, px, py, ox, oy);
  // the observer's position with its f32 rounding as the compensation term
  const float cx = float(double(s.obs_x) - s.obs_d[0]), cy = float(double(s.obs_y) - s.obs_d[1]),
              cz = float(double(s.obs_z) - s.obs_d[2]);
  init_ray<float>(sc, R, s.obs_x, s.obs_y, s.obs_z, cx, cy, cz, m.px, m.py, m.pz, m.E, -1.0f,
                  s.polar_side);
  R.h = -s.h_init;
  apply_start(R, m.start);
  if ((sc.options & kOptCertificate) && R.term == kTermActive && R.flags == 0u &&
      certify_capture(s, s.obs_x, s.obs_y, s.obs_z, R.px, R.py, R.pz, R.E, R.r, R.rdot)) {
    R.term = kTermHorizon;
    R.candidates = 2u;
    R.flags |= kFlagCertified;
  }
}

// The same at the scene's sample offset.
__device__ __forceinline__ void init_image_ray(const SceneDev& s, const SceneT<float>& sc,
                                               uint32_t px, uint32_t py, RayT<float>& R) {
  init_image_ray_at(s, sc, px, py, s.sub_x, s.sub_y, R);
}

__device__ void init_state_ray(const SceneDev& s, const SceneT<float>& sc, const double* y,
                               float dir, RayT<float>& R) {
  float x = float(y[1]), yy = float(y[2]), z = float(y[3]);
  float cx = float(double(x) - y[1]), cy = float(double(yy) - y[2]), cz = float(double(z) - y[3]);
  uint32_t side = y[3] > 0.0 ? 2u : (y[3] < 0.0 ? 0u : 1u);
  init_ray<float>(sc, R, x, yy, z, cx, cy, cz, float(y[5]), float(y[6]), float(y[7]),
                  float(-y[4]), dir, side);
  R.h = dir * s.h_init;
  apply_start(R, batch_start_test(s, y));
  // backward requests are pixel rays in all but their start: the same
  // certificate applies (certificate.h: its own domain checks refuse what it
  // cannot prove, and the ray is traced in full)
  if ((sc.options & kOptCertificate) && dir < 0.0f && R.term == kTermActive && R.flags == 0u &&
      certify_capture(s, x, yy, z, R.px, R.py, R.pz, R.E, R.r, R.rdot)) {
    R.term = kTermHorizon;
    R.candidates = 2u;
    R.flags |= kFlagCertified;
  }
}

// ---------------------------------------------------------------------------
// Surface radiance (once per surface hit): the intensity products and the
// RGBA16F range tests in f64, the band lookup's logarithms and powers of two
// in f32 (hardware log2 / exp2, about one ulp; a relative error near 1e-6 in
// the radiance, below the binary16 texel's resolution of 2^-11).

constexpr double kHalfMax = 65504.0;                        // largest finite binary16
constexpr double kMinNormal32 = 1.1754943508222875e-38;     // 2^-126

// log2 band fractions of a blackbody at log2 T = lt from the band table,
// linear in log2 T between entries (blackbody_surface_preview.wgsl:10-33);
// false outside the table or for a malformed entry.
__device__ __forceinline__ bool band_log2_fractions(const float* lut, float lt, float f[3]) {
  if (lut == nullptr || !(lt == lt)) return false;
  const float c = (lt - float(GL_BLACKBODY_LUT_MIN_LOG2_K)) * float(GL_BLACKBODY_LUT_PER_OCTAVE);
  if (!(c >= 0.0f && c <= float(GL_BLACKBODY_LUT_ENTRIES - 1))) return false;
  const int i = min(int(c), GL_BLACKBODY_LUT_ENTRIES - 2);
  const float w = c - float(i);
  const float4 lo = reinterpret_cast<const float4*>(lut)[i];
  const float4 hi = reinterpret_cast<const float4*>(lut)[i + 1];
  f[0] = lo.x + w * (hi.x - lo.x);
  f[1] = lo.y + w * (hi.y - lo.y);
  f[2] = lo.z + w * (hi.z - lo.z);
  return f[0] <= 0.0f && f[1] <= 0.0f && f[2] <= 0.0f;
}

// A positive radiance product resolved for RGBA16F: 0 below the f32 normal
// range, false above the largest half (surface_transport.wgsl:36-80).
__device__ __forceinline__ bool resolve_half(double v, double& out) {
  if (!(v <= kHalfMax)) return false;
  out = v < kMinNormal32 ? 0.0 : v;
  return true;
}

// v 2^f over f64's whole range: the power of two split as 2^n 2^(f - n),
// n = floor(f), its fraction from the f32 hardware exp2 and 2^n applied
// exactly (an f32 2^f alone would flush to zero below 2^-126, where
// v 2^f can still be a normal value).
__device__ __forceinline__ double scale_exp2(double v, float f) {
  if (!(f > -1.0e9f)) return 0.0;  // a zero band fraction (log2 = -inf)
  const float n = floorf(f);
  return ldexp(v * double(__builtin_amdgcn_exp2f(f - n)), int(n));
}

struct Rgb {
  float r, g, b;
};

// Observed radiance of a surface hit at source radius r (M = 1) with
// frequency ratio g: the inverse-cube emitter attenuated by the slab,
// I = I_6 (r/6)^-3 g^4 exp(-tau), plus the slab's own emission
// (surface_transport.wgsl:109-223); for a blackbody emitter per band, the
// incoming term scaled by the band fractions at the observed temperature
// g T_6 (r/6)^-3/4 and the slab term by those at its temperature
// (blackbody_surface_preview.wgsl:35-80). (-1, -1, -1) when the result is
// not representable in RGBA16F.
__device__ __noinline__ Rgb surface_radiance(const SceneDev& s, float r, float g) {
  const Rgb bad = {-1.0f, -1.0f, -1.0f};
  const double rr = double(r) / 6.0, gd = double(g);
  if (!(rr >= kMinNormal32) || !(gd > 0.0) || !(rr < 1.0e30)) return bad;
  const double g2 = gd * gd;
  const double incoming = s.intensity6_d * s.transmittance_d / (rr * rr * rr) * (g2 * g2);
  const double src = s.slab_emission_d;
  if (s.emission_model != GL_EMISSION_BLACKBODY) {
    double v;
    if (!resolve_half(incoming, v) || !(v + src <= kHalfMax)) return bad;
    const float o = float(v + src);
    return Rgb{o, o, o};
  }
  // log2 of the observed temperature, log2 T_6 - 3/4 log2(r/6) + log2 g
  const float lt = __builtin_amdgcn_logf(float(s.temperature6_d)) -
                   0.75f * __builtin_amdgcn_logf(float(rr)) + __builtin_amdgcn_logf(g);
  float fi[3], fs[3] = {0.0f, 0.0f, 0.0f};
  if (!band_log2_fractions(s.lut, lt, fi)) return bad;
  if (src > 0.0 &&
      !band_log2_fractions(s.lut, __builtin_amdgcn_logf(float(s.slab_temperature_d)), fs))
    return bad;
  double out[3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    double in_b, src_b = 0.0;
    if (!resolve_half(scale_exp2(incoming, fi[b]), in_b)) return bad;
    if (src > 0.0 && !resolve_half(scale_exp2(src, fs[b]), src_b)) return bad;
    if (!(in_b + src_b <= kHalfMax)) return bad;
    out[b] = in_b + src_b;
  }
  return Rgb{float(out[0]), float(out[1]), float(out[2])};
}

// Unit traversal direction dx/dlambda (times the trace direction) at an
// escape state (geodesic.rs:480-497).
struct Dir3 {
  float x, y, z;
  bool ok;
};

__device__ __forceinline__ Dir3 escape_direction(const SceneDev& s, float dir, float x, float y,
                                                 float z, float px, float py, float pz, float E) {
  const SceneT<float> sc = scene_f32(s);
  Deriv<float> d;
  Geom<float> g;
  if (!eval<float>(sc, x, y, z, px, py, pz, E, d, g)) return Dir3{0.0f, 0.0f, 0.0f, false};
  const float vx = dir * d.dx, vy = dir * d.dy, vz = dir * d.dz;
  const float inv = 1.0f / sqrtf(vx * vx + vy * vy + vz * vz);
  return Dir3{vx * inv, vy * inv, vz * inv, true};
}

// Source anchor (r, phi_s) and frequency ratio g of a surface hit, in f64
// from the localized f32 state and its Kahan compensation
// (surface.rs:75-146).
struct Anchor {
  float r, phi, g;
  bool ok;
};

__device__ __forceinline__ Anchor surface_anchor(const SceneDev& s, float x, float y, float z,
                                                 float cx, float cy, float cz, float px, float py,
                                                 float E) {
  const double xd = double(x) - double(cx), yd = double(y) - double(cy), zd = double(z) - double(cz);
  const Geom<double> g = geometry<double>(xd, yd, zd, double(s.mass), s.spin_d,
                                          s.spin_d * s.spin_d, s.charge2_d, double(s.chart));
  const double lz = xd * double(py) - yd * double(px);
  double o3[3];
  if (g.ok && surface_observable<double>(double(s.mass), s.spin_d, s.charge2_d, double(s.chart),
                                         double(s.omega), xd, yd, g.r, double(E), lz, o3))
    return Anchor{float(o3[0]), float(o3[1]), float(o3[2]), true};
  return Anchor{0.0f, 0.0f, 0.0f, false};
}

// Escape-map texel of a unit escape direction (equirectangular, nearest
// texel); black without a bound map.
__device__ __forceinline__ Rgb escape_texel(const MapView& map, const SceneDev& s, float dx, float dy,
                                            float dz) {
  if (!map.rgb || !(s.options & kOptEscapeMap)) return Rgb{0.0f, 0.0f, 0.0f};
  float lon = atan2f(dy, dx);
  float colat = acosf(fminf(fmaxf(dz, -1.0f), 1.0f));
  int iu = int((lon + 3.14159265f) * (0.15915494f * float(map.w)));
  int iv = int(colat * (0.31830989f * float(map.h)));
  iu = min(max(iu, 0), int(map.w) - 1);
  iv = min(max(iv, 0), int(map.h) - 1);
  size_t t = size_t(iv) * map.w + size_t(iu);
  size_t plane = size_t(map.w) * map.h;
  return Rgb{map.rgb[t], map.rgb[plane + t], map.rgb[2 * plane + t]};
}

// Failure colours of non-physical terminals (lensing_preview.wgsl:7-18).
__device__ __forceinline__ Rgb failure_colour(uint32_t term) {
  if (term == kTermSingularity) return Rgb{0.0f, 1.0f, 1.0f};
  if (term == kTermStepExhaustion) return Rgb{1.0f, 0.25f, 0.0f};
  if (term == kTermUncertain) return Rgb{1.0f, 1.0f, 0.0f};
  return Rgb{1.0f, 0.0f, 1.0f};
}

// One RGBA16F texel, round-to-nearest-even per channel, one 8-byte store.
__device__ __forceinline__ void store_rgba16f(uint16_t* plane, uint32_t idx, Rgb c, float a) {
  const _Float16 h[4] = {_Float16(c.r), _Float16(c.g), _Float16(c.b), _Float16(a)};
  uint2 v;
  __builtin_memcpy(&v, h, 8);
  reinterpret_cast<uint2*>(plane)[idx] = v;
}

// Where a ray started, to rebuild its exact f64 initial state at the end: a
// geodesic-batch request (its state) or a pixel at the scene's sample offset.
struct RaySource {
  const double* state8;  // batch request (8 doubles), or nullptr for a pixel
  uint32_t px, py;
};

__device__ __forceinline__ RaySource pixel_source(uint32_t px, uint32_t py) {
  return RaySource{nullptr, px, py};
}
__device__ __forceinline__ RaySource batch_source(const double* state8) {
  return RaySource{state8, 0u, 0u};
}
// The source of element idx of a frame slice: its batch request, or pixel
// (idx mod width, idx / width).
__device__ __forceinline__ RaySource frame_source(const FrameScene& f, const SceneDev& s,
                                                  uint32_t idx) {
  if (f.states) return batch_source(f.states + 8 * size_t(idx));
  return pixel_source(idx % s.width, idx / s.width);
}

// Near-critical rays handed to the f64 refine pass: (frame slice, element).
struct RefineRecord {
  uint32_t frame_scene, index;
  uint32_t bucket;  // the refine pass's claim-order bucket (refine_bucket)
};
struct RefineSink {
  RefineRecord* rec;  // nullptr: no refine pass in this launch
  unsigned int* count;
  uint32_t capacity;
  unsigned int* hist;  // records per claim bucket
};
__device__ __forceinline__ RefineSink refine_sink(const DeferBuf& df) {
  return RefineSink{df.refine, df.count + 2, df.refine_capacity, df.count + kRefineHist};
}
__device__ __forceinline__ RefineSink no_refine() {
  return RefineSink{nullptr, nullptr, 0u, nullptr};
}

__device__ __forceinline__ void source_state(const SceneDev& s, const double* state8, uint32_t px,
                                             uint32_t py, double y[8]) {
  if (state8) {
#pragma unroll
    for (int i = 0; i < 8; ++i) y[i] = state8[i];
  } else {
    image_state_f64(s, px, py, s.sub_x_d, s.sub_y_d, y);
  }
}

struct Verdict {
  double phi0;     // initial azimuth atan2(y, x)
  uint32_t extra;  // GL_STATUS_NEAR_CRITICAL, GL_STATUS_BRANCH_UNSETTLED or 0
};

// The near-critical test of a finished trace (near_critical.h), from the
// start test its trace carried (r_m, |mu| and the start bits, from the exact
// f64 initial state): a ray that came within 1.25 r_m (its smallest turning
// radius, or its terminal radius, r_near) is measured against its error
// budget, the largest drift it sampled at r from max(7/8 r_m, (r_m + r_+) / 2)
// to 2 r_m and, for an f32 trace, the rounding of its initial constants (the
// start bits): with
// |mu| <= near_critical_factor * budget the capture/scatter decision is
// unsettled (GL_STATUS_NEAR_CRITICAL), with |mu| <= branch_factor * budget
// the branch key and the time and azimuth spent near r_m are
// (GL_STATUS_BRANCH_UNSETTLED). phi0 is the azimuth of the initial position.
__device__ __noinline__ Verdict ray_verdict(const SceneDev& s, const double* state8,
                                            float r_near, float drift, uint32_t armed) {
  Verdict v;
  v.phi0 = state8 ? ::atan2(state8[2], state8[1]) : s.phi_observer_d;
  v.extra = 0u;
  const float r_m = decode_start_radius(armed);
  if (!(s.near_critical_factor > 0.0f) || !(r_m > 0.0f) || !(r_near <= 1.25f * r_m)) return v;
  const float mu = decode_start_margin(armed);
  float eps = drift;
  if (!(eps == eps)) eps = 1.0f;
  if (mu <= s.near_critical_factor * eps || (armed & kStartCritical))
    v.extra |= GL_STATUS_NEAR_CRITICAL;
  if (mu <= s.branch_factor * eps || (armed & kStartBranch))
    v.extra |= GL_STATUS_BRANCH_UNSETTLED;
  return v;
}

__device__ __forceinline__ bool event_class(uint32_t term) {
  return term == kTermHorizon || term == kTermEscape || term == kTermSurface ||
         term == kTermSingularity;
}

// Terminal observables from the localized state: the unit traversal
// direction at an escape (geodesic.rs:480-497), or the source anchor and
// frequency ratio of the prograde emitter in f64 at a surface hit
// (surface.rs:75-146); the event residual, the azimuth advance and the
// turning radius; then the packed status, the observable planes and the
// radiance. `term` is the class the trace reached (its observables are
// written), `status` the final status word (whose class may be Uncertain).
// Out of line (it runs once per ray and its f64 temporaries stay out of the
// trace loop's register allocation) and called with scalars only, so no ray
// state goes through memory.
__device__ __noinline__ void write_terminal(const OutPlanes& o, uint32_t idx, const SceneDev& s,
                                            MapView map, uint32_t term, uint32_t status, float dir,
                                            float x, float y, float z, float cx, float cy,
                                            float cz, float px, float py, float pz, float E,
                                            int32_t winding, double phi0, float r_turn) {
  const unsigned long long n = o.count;
  float ob[3] = {0.0f, 0.0f, 0.0f};
  if (term == kTermEscape) {
    const Dir3 v = escape_direction(s, dir, x, y, z, px, py, pz, E);
    ob[0] = v.x;
    ob[1] = v.y;
    ob[2] = v.z;
    if (!v.ok) status |= kFlagNonFinite << 8;
  } else if (term == kTermSurface) {
    const Anchor a = surface_anchor(s, x, y, z, cx, cy, cz, px, py, E);
    ob[0] = a.r;
    ob[1] = a.phi;
    ob[2] = a.g;
    if (!a.ok) status |= kFlagDenominator << 8;
  }
  const uint32_t shown = status & 15u;  // the class the texel shows
  if (o.status) o.status[idx] = status;
  if (o.observable) {
    o.observable[idx] = ob[0];
    o.observable[n + idx] = ob[1];
    o.observable[2 * n + idx] = ob[2];
  }
  if (o.event_residual || o.azimuth_advance) {
    const double xd = double(x) - double(cx), yd = double(y) - double(cy),
                 zd = double(z) - double(cz);
    if (o.event_residual) {
      // (event function at the terminal - its surface) / max(1, |surface|)
      // (geodesic_integration.wgsl:266-295)
      double res = 0.0;
      if (term == kTermSurface) {
        res = zd;
      } else if (term == kTermHorizon || term == kTermEscape || term == kTermSingularity) {
        const Geom<double> g = geometry<double>(xd, yd, zd, 1.0, s.spin_d, s.spin_d * s.spin_d,
                                                s.charge2_d, double(s.chart));
        if (term == kTermSingularity) {
          const double r2 = g.r * g.r;
          const double guard = double(s.guard);
          res = (r2 * r2 + s.spin_d * s.spin_d * zd * zd - guard) / fmax(1.0, guard);
        } else {
          const double target = term == kTermHorizon ? s.r_capture_d : double(s.r_escape);
          res = (g.r - target) / fmax(1.0, fabs(target));
        }
      }
      o.event_residual[idx] = float(res);
    }
    if (o.azimuth_advance) {
      // the sum of wrapped atan2 differences between committed states equals
      // the end azimuth minus the start azimuth plus 2 pi per branch-cut
      // crossing counted by the winding (geodesic.rs:406-418)
      const double adv = ::atan2(yd, xd) - phi0 + 6.283185307179586 * double(winding);
      o.azimuth_advance[idx] = float(adv);
    }
  }
  if (o.turning_radius) o.turning_radius[idx] = r_turn;
  if (o.radiance || o.rgba16f) {
    Rgb c = {0.0f, 0.0f, 0.0f};
    float alpha = -float(shown);
    if (shown == kTermEscape) {
      alpha = 1.0f;
      c = escape_texel(map, s, ob[0], ob[1], ob[2]);
    } else if (shown == kTermSurface) {
      c = ob[0] > 0.0f ? surface_radiance(s, ob[0], ob[2]) : Rgb{-1.0f, -1.0f, -1.0f};
      alpha = 2.0f;
    } else if (shown == kTermHorizon) {
      alpha = 0.0f;
    }
    if (o.radiance) {
      o.radiance[idx] = c.r;
      o.radiance[n + idx] = c.g;
      o.radiance[2 * n + idx] = c.b;
    }
    if (o.rgba16f) {
      if (shown == kTermSurface && c.r < 0.0f) {
        c = failure_colour(kTermFailure);
        alpha = -float(kTermFailure);
      } else if (shown != kTermEscape && shown != kTermSurface && shown != kTermHorizon) {
        c = failure_colour(shown);
      }
      store_rgba16f(o.rgba16f, idx, c, alpha);
    }
  }
}

// All output planes of one finished ray. The ray's class becomes Uncertain
// (the class it reached kept in status bits 16-19) when its terminal events
// tie, a turning lies inside the terminal's bracket, a drift exceeds the
// budget (geodesic_integration.wgsl:324-330), or it is near-critical
// (ray_verdict). A near-critical ray of a scene with the refine option goes
// to `sink` instead, unwritten, when there is room (returns false); the f64
// refine pass traces and writes it (refined = true: its exact f64 start
// adds no rounding to the budget, and it is not handed on again).
__device__ __forceinline__ bool write_ray(const RayT<float>& R, const OutPlanes& o, uint32_t idx,
                                          const SceneDev& s, const MapView& map, RaySource src,
                                          const RefineSink& sink, uint32_t frame_scene,
                                          bool refined = false) {
  const unsigned long long n = o.count;
  if (idx >= n) return true;
  const float r_near = R.r_turn >= 0.0f ? fminf(R.r_turn, R.r) : R.r;
  const Verdict v = ray_verdict(s, src.state8, r_near, R.d_crit, R.armed);
  const bool definite = event_class(R.term);
  const bool critical = definite && (v.extra & GL_STATUS_NEAR_CRITICAL);
  // the refine pass settles the branch as well as the class
  if (definite && (v.extra & (GL_STATUS_NEAR_CRITICAL | GL_STATUS_BRANCH_UNSETTLED)) && !refined &&
      (s.options & kOptRefine) && sink.rec) {
    const unsigned int slot = atomicAdd(sink.count, 1u);
    if (slot < sink.capacity) {
      const uint32_t b = refine_bucket(R.accepted + R.rejected);
      sink.rec[slot] = RefineRecord{frame_scene, idx, b};
      atomicAdd(sink.hist + b, 1u);
      return false;
    }
  }
  uint32_t cls = R.term, extra = (definite ? v.extra : 0u) | (refined ? GL_STATUS_REFINED : 0u);
  if (definite && (critical || (R.flags & (kFlagTied | kFlagTurningOrder | kFlagDrift)))) {
    cls = kTermUncertain;
    extra |= (R.term & 15u) << GL_STATUS_PROVISIONAL_SHIFT;
  }
  const uint32_t status = (cls & 15u) | ((R.candidates & 15u) << 4) | ((R.flags & 255u) << 8) | extra;
  if (o.steps) o.steps[idx] = R.accepted;
  if (o.rejected) o.rejected[idx] = R.rejected;
  if (o.travel_time) o.travel_time[idx] = R.dir * (R.t - R.ct);
  if (o.affine) o.affine[idx] = R.lam - R.clam;
  if (o.drift) {
    o.drift[idx] = R.d_null;
    o.drift[n + idx] = R.d_lz;
    o.drift[2 * n + idx] = R.d_q;
  }
  if (o.branch)
    o.branch[idx] = (min(R.turnings, 255u)) | (min(R.crossings, 255u) << 8) |
                    ((uint32_t(R.winding) & 255u) << 16) | ((R.polar_side & 3u) << 24);
  write_terminal(o, idx, s, map, R.term, status, R.dir, R.x, R.y, R.z, R.cx, R.cy, R.cz, R.px,
                 R.py, R.pz, R.E, R.winding, v.phi0, R.r_turn);
  return true;
}

// Scene constants of the trace loop held in VGPRs: as wave-uniform values
// they would occupy some twenty SGPRs next to the loop's exec masks, and the
// scalar file then spills to VGPR lanes with a readlane per use.
__device__ __forceinline__ void hold_in_vgprs(SceneT<float>& sc) {
  asm volatile("" : "+v"(sc.spin), "+v"(sc.spin2), "+v"(sc.charge2), "+v"(sc.chart));
  asm volatile("" : "+v"(sc.r_capture), "+v"(sc.r_defer), "+v"(sc.r_escape), "+v"(sc.disk_in),
               "+v"(sc.disk_out));
  asm volatile("" : "+v"(sc.band), "+v"(sc.rtol), "+v"(sc.atol), "+v"(sc.h_min), "+v"(sc.h_max),
               "+v"(sc.h_max_per_r), "+v"(sc.tie_tol), "+v"(sc.drift_limit), "+v"(sc.guard));
  asm volatile("" : "+v"(sc.safety), "+v"(sc.min_factor), "+v"(sc.max_factor));
}

// Null re-projection every reproject_interval-th DP5 attempt of a ray
// (attempt = 1, 2, ...: accepted and rejected attempts both count), so a
// ray's arithmetic does not depend on which kernel or wavefront traces it.
__device__ __forceinline__ bool reproject_now(const SceneT<float>& sc, uint32_t attempt) {
  return sc.reproject_interval != 0u && (attempt % sc.reproject_interval) == 0u;
}

// Invariant drifts are sampled every fourth attempt of a ray (and at every
// re-projection and termination).
__device__ __forceinline__ bool drift_now(uint32_t attempt) { return (attempt & 3u) == 0u; }

__device__ __forceinline__ bool wants_defer(const SceneT<float>& sc, const RayT<float>& R) {
  return (sc.options & kOptF64Horizon) && R.r < sc.r_defer;
}

__device__ __forceinline__ bool try_defer(const DeferBuf& df, const RayT<float>& R, uint32_t fs,
                                          uint32_t idx) {
  if (df.rec == nullptr) return false;
  unsigned int slot = atomicAdd(df.count, 1u);
  if (slot >= df.capacity) return false;
  df.rec[slot].ray = R;
  df.rec[slot].frame_scene = fs;
  df.rec[slot].index = idx;
  const uint32_t b = defer_bucket(decode_start_radius(R.armed));
  df.rec[slot].bucket = b;
  atomicAdd(df.count + kDeferHist + b, 1u);
  return true;
}

// The f32 stepping loop of one ray from attempt `it`: DP5 attempts until
// the ray ends, or until it enters the f64 island and its record goes to the
// finisher (deferred = true). A deferral the buffer cannot take (it is full
// for the rest of the launch) leaves the ray in f32. The loop exits when no
// lane of the wavefront has a ray left to step, so its exit is
// wavefront-uniform: a lane whose ray is done idles, masked, until then. A
// loop that lanes leave at different trips makes the compiler copy every
// loop-carried value of the ray state on each trip (its value at the lane's
// exit must survive the later trips of the other lanes).
__device__ __forceinline__ void trace_f32(const SceneT<float>& sc, RayT<float>& R, float omega,
                                          const DeferBuf& df, uint32_t fs, uint32_t idx,
                                          uint32_t& it, bool& deferred) {
  bool may_defer = (sc.options & kOptF64Horizon) != 0u;
  // the re-projection clock, it mod reproject_interval, kept by increments
  // (reproject_now's test without a division per attempt)
  const uint32_t period = sc.reproject_interval;
  uint32_t phase = period != 0u ? it % period : 1u;
  for (;;) {
    const bool live = !deferred && R.term == kTermActive;
    if (!__any(live)) break;
    if (live) {
      if (!(may_defer && R.r < sc.r_defer)) {
        step<float>(sc, R, omega, period != 0u && phase == 0u, drift_now(it));
        ++it;
        phase = phase + 1u == period ? 0u : phase + 1u;
      } else if (try_defer(df, R, fs, idx)) {
        deferred = true;
      } else {
        may_defer = false;
      }
    }
  }
}

}  // namespace gl
