// Device-side pieces shared by the trace kernels and the adjoint kernels:
// the escape-map view, the deferred-ray buffer, pixel mapping, the initial
// ray, and the output writer.
#pragma once

#include <hip/hip_runtime.h>

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

__device__ __noinline__ Momentum image_momentum(const SceneDev& s, uint32_t px, uint32_t py) {
  const SceneT<float> sc = scene_f32(s);
  float nx = 2.0f * (float(px) + s.sub_x) / float(s.width) - 1.0f;
  float ny = 1.0f - 2.0f * (float(py) + s.sub_y) / float(s.height);
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

__device__ __forceinline__ void init_image_ray(const SceneDev& s, const SceneT<float>& sc,
                                               uint32_t px, uint32_t py, RayT<float>& R) {
  Momentum m = image_momentum(s, px, py);
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

__device__ void init_state_ray(const SceneDev& s, const SceneT<float>& sc, const double* y,
                               float dir, RayT<float>& R) {
  float x = float(y[1]), yy = float(y[2]), z = float(y[3]);
  float cx = float(double(x) - y[1]), cy = float(double(yy) - y[2]), cz = float(double(z) - y[3]);
  uint32_t side = y[3] > 0.0 ? 2u : (y[3] < 0.0 ? 0u : 1u);
  init_ray<float>(sc, R, x, yy, z, cx, cy, cz, float(y[5]), float(y[6]), float(y[7]),
                  float(-y[4]), dir, side);
  R.h = dir * s.h_init;
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
    const SceneT<float> sc = scene_f32(s);
    Deriv<float> d;
    Geom<float> g;
    if (eval<float>(sc, x, y, z, px, py, pz, E, d, g)) {
      float vx = dir * d.dx, vy = dir * d.dy, vz = dir * d.dz;
      float inv = 1.0f / sqrtf(vx * vx + vy * vy + vz * vz);
      ob[0] = vx * inv;
      ob[1] = vy * inv;
      ob[2] = vz * inv;
    } else {
      flags |= kFlagNonFinite;
    }
  } else if (term == kTermSurface) {
    double xd = double(x) - double(cx), yd = double(y) - double(cy), zd = double(z) - double(cz);
    Geom<double> g = geometry<double>(xd, yd, zd, double(s.mass), s.spin_d, s.spin_d * s.spin_d,
                                      s.charge2_d, double(s.chart));
    double lz = xd * double(py) - yd * double(px);
    double o3[3];
    if (g.ok && surface_observable<double>(double(s.mass), s.spin_d, s.charge2_d, double(s.chart),
                                           double(s.omega), xd, yd, g.r, double(E), lz, o3)) {
      ob[0] = float(o3[0]);
      ob[1] = float(o3[1]);
      ob[2] = float(o3[2]);
    } else {
      flags |= kFlagDenominator;
    }
  }
  if (o.status) o.status[idx] = (term & 15u) | ((candidates & 15u) << 4) | ((flags & 255u) << 8);
  if (o.observable) {
    o.observable[idx] = ob[0];
    o.observable[n + idx] = ob[1];
    o.observable[2 * n + idx] = ob[2];
  }
  if (o.radiance) {
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    if (term == kTermEscape && map.rgb && (s.options & kOptEscapeMap)) {
      // equirectangular lookup of the escape direction
      float lon = atan2f(ob[1], ob[0]);
      float colat = acosf(fminf(fmaxf(ob[2], -1.0f), 1.0f));
      int iu = int((lon + 3.14159265f) * (0.15915494f * float(map.w)));
      int iv = int(colat * (0.31830989f * float(map.h)));
      iu = min(max(iu, 0), int(map.w) - 1);
      iv = min(max(iv, 0), int(map.h) - 1);
      size_t t = size_t(iv) * map.w + size_t(iu);
      size_t plane = size_t(map.w) * map.h;
      c0 = map.rgb[t];
      c1 = map.rgb[plane + t];
      c2 = map.rgb[2 * plane + t];
    } else if (term == kTermSurface && ob[0] > 0.0f) {
      float r = ob[0], g = ob[2];
      float ratio = r / s.disk_in;
      c0 = g * s.temperature_inner * powf(ratio, -0.75f);  // observed temperature
      float rr = r / 6.0f;
      float g2 = g * g;
      c1 = g2 * g2 / (rr * rr * rr);  // g^4 I_6 (r/6)^-3 with I_6 = 1
    }
    o.radiance[idx] = c0;
    o.radiance[n + idx] = c1;
    o.radiance[2 * n + idx] = c2;
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
