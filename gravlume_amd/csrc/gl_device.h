// Device-side pieces shared by the trace kernels and the adjoint kernels:
// the escape-map view, the deferred-ray buffer, pixel mapping, the initial
// ray, and the output writer.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gravlume.h"
#include "certificate.h"
#include "gl_types.h"
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
};

struct DeferBuf {
  DeferRecord* rec;
  unsigned int* count;
  uint32_t capacity;
};

// ---------------------------------------------------------------------------
// Device helpers

__device__ __forceinline__ void image_pixel(uint32_t item, uint32_t tiles_x, uint32_t& px,
                                            uint32_t& py) {
  uint32_t tile = item >> 6, t = item & 63u;
  px = (tile % tiles_x) * 8u + (t & 7u);
  py = (tile / tiles_x) * 8u + (t >> 3);
}

// Sight momentum of pixel (px, py): direction from the tetrad, future-directed
// p = omega (u + n), lowered with the Kerr-Schild metric at the observer
// (scene.rs:211-271, view.rs:107-123), then projected onto the null cone in
// f64. Out of line and returned in registers (four dwords), so the ray state
// of the calling kernel never has its address taken and stays in VGPRs.
struct Momentum {
  float px, py, pz, E;
};

__device__ __noinline__ Momentum image_momentum(const SceneDev& s, uint32_t px, uint32_t py,
                                                float ox, float oy) {
  const SceneT<float> sc = scene_f32(s);
  float nx = 2.0f * (float(px) + ox) / float(s.width) - 1.0f;
  float ny = 1.0f - 2.0f * (float(py) + oy) / float(s.height);
  float sx = s.aspect * s.tan_half_fov * nx;
  float sy = s.tan_half_fov * ny;
  float inv = 1.0f / sqrtf(1.0f + sx * sx + sy * sy);
  float d[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) d[i] = (sx * s.right[i] + sy * s.up[i] - s.arrival[i]) * inv;
  // p^mu = omega (u - d), u = (u_t, 0, 0, 0)
  float pu[4] = {s.omega * (s.u_t - d[0]), -s.omega * d[1], -s.omega * d[2], -s.omega * d[3]};
  Geom<float> g = geometry<float>(s.obs_x, s.obs_y, s.obs_z, sc.mass, sc.spin, sc.spin2,
                                  sc.charge2, sc.chart);
  float lp = pu[0] + g.lx * pu[1] + g.ly * pu[2] + g.lz * pu[3];
  float fl = g.f * lp;
  float p_t = -pu[0] + fl;
  Momentum m;
  m.px = pu[1] + fl * g.lx;
  m.py = pu[2] + fl * g.ly;
  m.pz = pu[3] + fl * g.lz;
  m.E = -p_t;
  RayT<float> at;  // the observer event, for the projection
  at.x = s.obs_x;
  at.y = s.obs_y;
  at.z = s.obs_z;
  at.cx = at.cy = at.cz = 0.0f;
  at.E = m.E;
  reproject_null<float>(sc, at, m.px, m.py, m.pz);
  return m;
}

// The interval radial-capture certificate at the observer (certificate.h),
// from the ray's constants of motion there. Out of line, once per ray.
__device__ __noinline__ bool certify_capture(const SceneDev& s, float px, float py, float pz,
                                             float E, float r_obs, float dr_future) {
  const Geom<float> g = geometry<float>(s.obs_x, s.obs_y, s.obs_z, s.mass, s.spin, s.spin2,
                                        s.charge2, s.chart);
  const float L = s.obs_x * py - s.obs_y * px;
  const float Q = carter<float>(g, s.obs_x, s.obs_y, s.obs_z, px, py, pz, E, s.spin2);
  const float q = float(sqrt(s.charge2_d));
  return capture_certified(s.mass, s.spin, q, s.r_horizon, r_obs,
                           s.obs_x * s.obs_x + s.obs_y * s.obs_y, dr_future, E, L, Q);
}

// The backward ray through point (px + ox, py + oy) of the viewport, with
// the capture certificate applied when the scene enables it.
__device__ __forceinline__ void init_image_ray_at(const SceneDev& s, const SceneT<float>& sc,
                                                  uint32_t px, uint32_t py, float ox, float oy,
                                                  RayT<float>& R) {
  Momentum m = image_momentum(s, px, py, ox, oy);
  init_ray<float>(sc, R, s.obs_x, s.obs_y, s.obs_z, 0.0f, 0.0f, 0.0f, m.px, m.py, m.pz, m.E,
                  -1.0f, s.polar_side);
  R.h = -s.h_init;
  if ((sc.options & kOptCertificate) && R.term == kTermActive && R.flags == 0u &&
      certify_capture(s, R.px, R.py, R.pz, R.E, R.r, R.rdot)) {
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
}

// ---------------------------------------------------------------------------
// Surface radiance (f64, once per surface hit)

constexpr double kHalfMax = 65504.0;                        // largest finite binary16
constexpr double kMinNormal32 = 1.1754943508222875e-38;     // 2^-126

// log2 band fractions of a blackbody at T from the band table, linear in
// log2 T between entries (blackbody_surface_preview.wgsl:10-33); false
// outside the table or for a malformed entry.
__device__ __forceinline__ bool band_log2_fractions(const float* lut, double t, double f[3]) {
  if (lut == nullptr || !(t > 0.0) || !(t < 1.0e300)) return false;
  const double c = (::log2(t) - GL_BLACKBODY_LUT_MIN_LOG2_K) * GL_BLACKBODY_LUT_PER_OCTAVE;
  if (!(c >= 0.0 && c <= double(GL_BLACKBODY_LUT_ENTRIES - 1))) return false;
  const int i = min(int(c), GL_BLACKBODY_LUT_ENTRIES - 2);
  const double w = c - double(i);
  const float4 lo = reinterpret_cast<const float4*>(lut)[i];
  const float4 hi = reinterpret_cast<const float4*>(lut)[i + 1];
  f[0] = double(lo.x) + w * (double(hi.x) - double(lo.x));
  f[1] = double(lo.y) + w * (double(hi.y) - double(lo.y));
  f[2] = double(lo.z) + w * (double(hi.z) - double(lo.z));
  return f[0] <= 0.0 && f[1] <= 0.0 && f[2] <= 0.0;
}

// A positive radiance product resolved for RGBA16F: 0 below the f32 normal
// range, false above the largest half (surface_transport.wgsl:36-80).
__device__ __forceinline__ bool resolve_half(double v, double& out) {
  if (!(v <= kHalfMax)) return false;
  out = v < kMinNormal32 ? 0.0 : v;
  return true;
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
  double fi[3], fs[3] = {0.0, 0.0, 0.0};
  const double t_obs = s.temperature6_d / __builtin_sqrt(rr * __builtin_sqrt(rr)) * gd;
  if (!band_log2_fractions(s.lut, t_obs, fi)) return bad;
  if (src > 0.0 && !band_log2_fractions(s.lut, s.slab_temperature_d, fs)) return bad;
  double out[3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    double in_b, src_b = 0.0;
    if (!resolve_half(incoming * ::exp2(fi[b]), in_b)) return bad;
    if (src > 0.0 && !resolve_half(src * ::exp2(fs[b]), src_b)) return bad;
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

// Terminal observables from the localized state: the unit traversal
// direction at an escape (geodesic.rs:480-497), or the source anchor and
// frequency ratio of the prograde emitter in f64 at a surface hit
// (surface.rs:75-146); then the packed status, the observable planes and the
// radiance. Out of line (it runs once per ray and its f64 temporaries stay
// out of the trace loop's register allocation) and called with scalars only,
// so no ray state goes through memory.
__device__ __noinline__ void write_terminal(const OutPlanes& o, uint32_t idx, const SceneDev& s,
                                            MapView map, uint32_t term, uint32_t candidates,
                                            uint32_t flags, float dir, float x, float y, float z,
                                            float cx, float cy, float cz, float px, float py,
                                            float pz, float E) {
  const unsigned long long n = o.count;
  float ob[3] = {0.0f, 0.0f, 0.0f};
  if (term == kTermEscape) {
    const Dir3 v = escape_direction(s, dir, x, y, z, px, py, pz, E);
    ob[0] = v.x;
    ob[1] = v.y;
    ob[2] = v.z;
    if (!v.ok) flags |= kFlagNonFinite;
  } else if (term == kTermSurface) {
    const Anchor a = surface_anchor(s, x, y, z, cx, cy, cz, px, py, E);
    ob[0] = a.r;
    ob[1] = a.phi;
    ob[2] = a.g;
    if (!a.ok) flags |= kFlagDenominator;
  }
  if (o.status) o.status[idx] = (term & 15u) | ((candidates & 15u) << 4) | ((flags & 255u) << 8);
  if (o.observable) {
    o.observable[idx] = ob[0];
    o.observable[n + idx] = ob[1];
    o.observable[2 * n + idx] = ob[2];
  }
  if (o.radiance || o.rgba16f) {
    Rgb c = {0.0f, 0.0f, 0.0f};
    float alpha = -float(term);
    if (term == kTermEscape) {
      alpha = 1.0f;
      c = escape_texel(map, s, ob[0], ob[1], ob[2]);
    } else if (term == kTermSurface) {
      c = ob[0] > 0.0f ? surface_radiance(s, ob[0], ob[2]) : Rgb{-1.0f, -1.0f, -1.0f};
      alpha = 2.0f;
    } else if (term == kTermHorizon) {
      alpha = 0.0f;
    }
    if (o.radiance) {
      o.radiance[idx] = c.r;
      o.radiance[n + idx] = c.g;
      o.radiance[2 * n + idx] = c.b;
    }
    if (o.rgba16f) {
      if (term == kTermSurface && c.r < 0.0f) {
        c = failure_colour(kTermFailure);
        alpha = -float(kTermFailure);
      } else if (term != kTermEscape && term != kTermSurface && term != kTermHorizon) {
        c = failure_colour(term);
      }
      store_rgba16f(o.rgba16f, idx, c, alpha);
    }
  }
}

// All output planes of one finished ray.
__device__ __forceinline__ void write_ray(const RayT<float>& R, const OutPlanes& o, uint32_t idx,
                                          const SceneDev& s, const MapView& map) {
  const unsigned long long n = o.count;
  if (idx >= n) return;
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
  write_terminal(o, idx, s, map, R.term, R.candidates, R.flags, R.dir, R.x, R.y, R.z, R.cx, R.cy,
                 R.cz, R.px, R.py, R.pz, R.E);
}

// Scene constants of the trace loop held in VGPRs: as wave-uniform values
// they would occupy some twenty SGPRs next to the loop's exec masks, and the
// scalar file then spills to VGPR lanes with a readlane per use.
__device__ __forceinline__ void hold_in_vgprs(SceneT<float>& sc) {
  asm volatile("" : "+v"(sc.spin), "+v"(sc.spin2), "+v"(sc.charge2), "+v"(sc.chart));
  asm volatile("" : "+v"(sc.r_capture), "+v"(sc.r_horizon), "+v"(sc.r_escape), "+v"(sc.disk_in),
               "+v"(sc.disk_out));
  asm volatile("" : "+v"(sc.band), "+v"(sc.rtol), "+v"(sc.atol), "+v"(sc.h_min), "+v"(sc.h_max),
               "+v"(sc.h_max_per_r), "+v"(sc.tie_tol), "+v"(sc.drift_limit), "+v"(sc.guard));
  asm volatile("" : "+v"(sc.safety), "+v"(sc.min_factor), "+v"(sc.max_factor));
}

// Null re-projection on a schedule every lane of a wavefront shares: the
// f64 branch runs for the whole wave at once instead of once per lane.
__device__ __forceinline__ bool reproject_now(const SceneT<float>& sc, uint32_t wave_step) {
  return sc.reproject_interval != 0u && (wave_step % sc.reproject_interval) == 0u;
}

// Invariant drifts are sampled every fourth attempt (and at every
// re-projection and termination), on the same wavefront-uniform clock.
__device__ __forceinline__ bool drift_now(uint32_t wave_step) { return (wave_step & 3u) == 0u; }

__device__ __forceinline__ bool wants_defer(const SceneT<float>& sc, const RayT<float>& R) {
  return (sc.options & kOptF64Horizon) && sc.r_horizon > 0.0f && R.r < 1.05f * sc.r_horizon;
}

__device__ __forceinline__ bool try_defer(const DeferBuf& df, const RayT<float>& R, uint32_t fs,
                                          uint32_t idx) {
  if (df.rec == nullptr) return false;
  unsigned int slot = atomicAdd(df.count, 1u);
  if (slot >= df.capacity) return false;
  df.rec[slot].ray = R;
  df.rec[slot].frame_scene = fs;
  df.rec[slot].index = idx;
  return true;
}

}  // namespace gl
