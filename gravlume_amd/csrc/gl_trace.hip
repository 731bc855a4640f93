// HIP kernels and C ABI of the gfx950 Kerr-Newman null-geodesic trace.
//
// Kernels (one ray per lane, wave64):
//   k_plain_image      one 8x8 tile per 64-lane workgroup over a tile region
//                      (the reference's encode_batch dispatch, trace.rs:756-784)
//   k_persistent       resident workgroups pull rays of many scenes from a
//                      device work queue; each wavefront refills its idle
//                      lanes with one atomic per refill (ballot + mbcnt)
//   k_geodesic_batch   arbitrary f64 initial states (GeodesicBatch::trace_ordered)
//   k_finish_f64       rays that entered r < 1.05 r_+ finish with f64 DP5 steps
//   k_queue_reset      resets the work queue (a kernel, so graph replays work)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gravlume.h"
#include "gl_scene.h"
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

__device__ __forceinline__ void init_image_ray(const SceneDev& s, const SceneT<float>& sc,
                                               uint32_t px, uint32_t py, RayT<float>& R) {
  Momentum m = image_momentum(s, px, py);
  init_ray<float>(sc, R, s.obs_x, s.obs_y, s.obs_z, 0.0f, 0.0f, 0.0f, m.px, m.py, m.pz, m.E,
                  -1.0f, s.polar_side);
  R.h = -s.h_init;
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
  asm volatile("" : "+v"(sc.mass), "+v"(sc.spin), "+v"(sc.spin2), "+v"(sc.charge2), "+v"(sc.chart));
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

// ---------------------------------------------------------------------------
// Kernels

__global__ void __launch_bounds__(64) k_plain_image(const SceneDev* __restrict__ scenes,
                                                    const FrameScene* __restrict__ frame,
                                                    uint32_t tile_x0, uint32_t tile_y0,
                                                    uint32_t tiles_x, MapView map, DeferBuf df) {
  const FrameScene& fsc = frame[0];
  const SceneDev& s = scenes[fsc.scene];
  SceneT<float> sc = scene_f32(s);
  hold_in_vgprs(sc);
  uint32_t tx = tile_x0 + blockIdx.x % tiles_x;
  uint32_t ty = tile_y0 + blockIdx.x / tiles_x;
  uint32_t px = tx * 8u + (threadIdx.x & 7u);
  uint32_t py = ty * 8u + (threadIdx.x >> 3);
  if (px >= s.width || py >= s.height) return;
  uint32_t idx = py * s.width + px;
  RayT<float> R;
  init_image_ray(s, sc, px, py, R);
  for (uint32_t it = 1; R.term == kTermActive; ++it) {
    if (wants_defer(sc, R) && try_defer(df, R, 0u, idx)) return;
    step<float>(sc, R, s.omega, reproject_now(sc, it), drift_now(it));
  }
  write_ray(R, fsc.out, idx, s, map);
}

__global__ void __launch_bounds__(64) k_geodesic_batch(const SceneDev* __restrict__ scenes,
                                                       const FrameScene* __restrict__ frame,
                                                       const double* __restrict__ states,
                                                       const float* __restrict__ dirs,
                                                       uint32_t n, MapView map, DeferBuf df) {
  const FrameScene& fsc = frame[0];
  const SceneDev& s = scenes[fsc.scene];
  const SceneT<float> sc = scene_f32(s);
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  RayT<float> R;
  init_state_ray(s, sc, states + 8 * size_t(i), dirs[i] >= 0.0f ? 1.0f : -1.0f, R);
  for (uint32_t it = 1; R.term == kTermActive; ++it) {
    if (wants_defer(sc, R) && try_defer(df, R, 0u, i)) return;
    step<float>(sc, R, s.omega, reproject_now(sc, it), drift_now(it));
  }
  write_ray(R, fsc.out, i, s, map);
}

__device__ __forceinline__ uint32_t find_frame_scene(const FrameScene* frame, uint32_t n,
                                                     unsigned long long item) {
  uint32_t lo = 0, hi = n;  // largest i with first_item <= item
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (frame[mid].first_item <= item) lo = mid; else hi = mid;
  }
  return lo;
}

// kWaves: minimum waves per SIMD the register allocation must allow
// (the second __launch_bounds__ argument); the autotuner picks it.
template <bool kSingle, int kWaves>
__global__ void __launch_bounds__(256, kWaves) k_persistent(const SceneDev* __restrict__ scenes,
                                                    const FrameScene* __restrict__ frame,
                                                    uint32_t n_frame, QueueDev* queue, MapView map,
                                                    DeferBuf df, uint32_t interval,
                                                    uint32_t threshold) {
  const unsigned long long total = queue->total;
  const uint32_t lane = __lane_id();
  RayT<float> R;
  R.term = kTermActive;
  bool active = false;
  uint32_t fs = 0, idx = 0;
  SceneT<float> sc;
  float omega = 0.0f;
  if (kSingle) {
    sc = scene_f32(scenes[frame[0].scene]);
    omega = scenes[frame[0].scene].omega;
    hold_in_vgprs(sc);
  }
  uint32_t lane_attempts = 0;  // per lane: DP5 attempts made
  uint32_t refills = 0;
  uint32_t wave_steps = 0;  // wave-uniform attempt counter (lane slots / 64)
  bool drained = false;  // wave-uniform: the queue has no items left
  for (;;) {
    if (!drained) {
      unsigned long long free_mask = __ballot(!active);
      uint32_t nfree = __popcll(free_mask);
      if (nfree >= threshold || nfree == 64u) {
        uint32_t leader = __ffsll((long long)free_mask) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(&queue->next, (unsigned long long)nfree);
        base = __shfl(base, leader);
        refills += 1;
        if (base + nfree >= total) drained = true;
        if (!active) {
          uint32_t rank = __popcll(free_mask & ((1ull << lane) - 1ull));
          unsigned long long item = base + rank;
          if (item < total) {
            fs = kSingle ? 0u : find_frame_scene(frame, n_frame, item);
            const FrameScene& f = frame[fs];
            const SceneDev& s = scenes[f.scene];
            if (!kSingle) {
              sc = scene_f32(s);
              omega = s.omega;
            }
            uint32_t local = uint32_t(item - f.first_item);
            uint32_t px, py;
            image_pixel(local, s.tiles_x, px, py);
            if (px < s.width && py < s.height) {
              idx = py * s.width + px;
              init_image_ray(s, sc, px, py, R);
              active = true;
              if (R.term != kTermActive) {
                write_ray(R, f.out, idx, s, map);
                active = false;
              }
            }
          }
        }
      }
    }
    if (__ballot(active) == 0ull) {
      if (drained) break;
      continue;
    }
    for (uint32_t it = 0; it < interval; ++it) {
      ++wave_steps;
      if (active) {
        ++lane_attempts;
        if (wants_defer(sc, R) && try_defer(df, R, fs, idx)) {
          active = false;
        } else {
          step<float>(sc, R, omega, reproject_now(sc, wave_steps), drift_now(wave_steps));
          if (R.term != kTermActive) {
            const FrameScene& f = frame[fs];
            write_ray(R, f.out, idx, scenes[f.scene], map);
            active = false;
          }
        }
      }
      if (!drained && __popcll(__ballot(!active)) >= threshold) break;
      if (drained && __ballot(active) == 0ull) break;
    }
  }
  for (int off = 32; off > 0; off >>= 1) lane_attempts += __shfl_xor(lane_attempts, off);
  if (lane == 0) {
    atomicAdd(&queue->attempts, (unsigned long long)lane_attempts);
    atomicAdd(&queue->slots, 64ull * wave_steps);
    atomicAdd(&queue->refills, (unsigned long long)refills);
  }
}

// Whole-tile mode of the persistent kernel (refill threshold 64): each
// wavefront takes one 8x8 tile (64 queue items, all of one scene) at a time
// and traces it to completion before taking the next, so neighbouring pixels
// stay together in a wave as in the plain dispatch, while the workgroups stay
// resident and pull tiles of every scene from the one queue.
template <bool kSingle, int kWaves>
__global__ void __launch_bounds__(256, kWaves) k_persistent_tiles(
    const SceneDev* __restrict__ scenes, const FrameScene* __restrict__ frame, uint32_t n_frame,
    QueueDev* queue, MapView map, DeferBuf df, uint32_t /*interval*/, uint32_t /*threshold*/) {
  const unsigned long long total = queue->total;
  const uint32_t lane = __lane_id();
  uint32_t lane_attempts = 0, tiles = 0, wave_steps = 0;
  for (;;) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&queue->next, 64ull);
    base = __shfl(base, 0);
    if (base >= total) break;
    ++tiles;
    const unsigned long long item = base + lane;
    const uint32_t fs = kSingle ? 0u : find_frame_scene(frame, n_frame, item);
    const FrameScene& f = frame[fs];
    const SceneDev& s = scenes[f.scene];
    SceneT<float> sc = scene_f32(s);
    hold_in_vgprs(sc);
    uint32_t px, py;
    image_pixel(uint32_t(item - f.first_item), s.tiles_x, px, py);
    uint32_t it = 0;
    if (px < s.width && py < s.height) {
      const uint32_t idx = py * s.width + px;
      RayT<float> R;
      init_image_ray(s, sc, px, py, R);
      bool deferred = false;
      for (it = 1; R.term == kTermActive; ++it) {
        if (wants_defer(sc, R) && try_defer(df, R, fs, idx)) {
          deferred = true;
          break;
        }
        step<float>(sc, R, s.omega, reproject_now(sc, it), drift_now(it));
      }
      lane_attempts += it - 1;
      if (!deferred) write_ray(R, f.out, idx, s, map);
    }
    uint32_t longest = it;  // wave iterations of this tile: its longest lane
    for (int off = 32; off > 0; off >>= 1) longest = max(longest, (uint32_t)__shfl_xor(longest, off));
    wave_steps += longest > 0u ? longest - 1u : 0u;
  }
  for (int off = 32; off > 0; off >>= 1) lane_attempts += __shfl_xor(lane_attempts, off);
  if (lane == 0) {
    atomicAdd(&queue->attempts, (unsigned long long)lane_attempts);
    atomicAdd(&queue->slots, 64ull * wave_steps);
    atomicAdd(&queue->refills, (unsigned long long)tiles);
  }
}

__global__ void __launch_bounds__(64) k_finish_f64(const SceneDev* __restrict__ scenes,
                                                   const FrameScene* __restrict__ frame,
                                                   MapView map, DeferBuf df) {
  unsigned int n = *df.count;
  if (n > df.capacity) n = df.capacity;
  for (unsigned int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const DeferRecord& rec = df.rec[i];
    const FrameScene& f = frame[rec.frame_scene];
    const SceneDev& s = scenes[f.scene];
    const SceneT<double> sd = scene_f64(s);
    const RayT<float>& Rf = rec.ray;
    RayT<double> R;
    R.x = double(Rf.x) - double(Rf.cx);
    R.y = double(Rf.y) - double(Rf.cy);
    R.z = double(Rf.z) - double(Rf.cz);
    R.cx = R.cy = R.cz = 0.0;
    R.px = Rf.px;
    R.py = Rf.py;
    R.pz = Rf.pz;
    R.E = Rf.E;
    R.h = Rf.h;
    R.t = double(Rf.t) - double(Rf.ct);
    R.ct = 0.0;
    R.lam = double(Rf.lam) - double(Rf.clam);
    R.clam = 0.0;
    R.lz0 = Rf.lz0;
    R.q0 = Rf.q0;
    R.d_null = Rf.d_null;
    R.d_lz = Rf.d_lz;
    R.d_q = Rf.d_q;
    R.accepted = Rf.accepted;
    R.rejected = Rf.rejected;
    R.consec = Rf.consec;
    R.turnings = Rf.turnings;
    R.crossings = Rf.crossings;
    R.winding = Rf.winding;
    R.armed = Rf.armed;
    R.term = Rf.term;
    R.candidates = Rf.candidates;
    R.flags = Rf.flags;
    R.polar_side = Rf.polar_side;
    R.dir = Rf.dir;
    if (!refresh_derivative<double>(sd, R)) {
      R.term = kTermFailure;
      R.flags |= kFlagNonFinite;
    }
    for (uint32_t it = 1; R.term == kTermActive; ++it)
      step<double>(sd, R, double(s.omega), false, drift_now(it));
    RayT<float> W;
    W.x = float(R.x);
    W.y = float(R.y);
    W.z = float(R.z);
    W.cx = float(double(W.x) - (R.x - R.cx));
    W.cy = float(double(W.y) - (R.y - R.cy));
    W.cz = float(double(W.z) - (R.z - R.cz));
    W.px = float(R.px);
    W.py = float(R.py);
    W.pz = float(R.pz);
    W.E = float(R.E);
    W.term = R.term;
    W.candidates = R.candidates;
    W.flags = R.flags;
    W.accepted = R.accepted;
    W.rejected = R.rejected;
    W.t = float(R.t);
    W.ct = 0.0f;
    W.lam = float(R.lam);
    W.clam = 0.0f;
    W.dir = R.dir;
    W.d_null = float(R.d_null);
    W.d_lz = float(R.d_lz);
    W.d_q = float(R.d_q);
    W.turnings = R.turnings;
    W.crossings = R.crossings;
    W.winding = R.winding;
    W.polar_side = R.polar_side;
    write_ray(W, f.out, rec.index, s, map);
  }
}

__global__ void k_queue_reset(QueueDev* q, unsigned long long total, unsigned int* defer_count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    q->next = 0;
    q->total = total;
    q->refills = 0;
    q->attempts = 0;
    q->slots = 0;
    if (defer_count) *defer_count = 0;
  }
}

}  // namespace gl

// ===========================================================================
// Host side

using namespace gl;

namespace {

thread_local std::string g_last_error;

gl_status fail(gl_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

gl_status hip_fail(hipError_t e, const char* what) {
  return fail(GL_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define GL_HIP(call)                                   \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return hip_fail(e_, #call);  \
  } while (0)

OutPlanes to_planes(const gl_ray_outputs& o) {
  OutPlanes p;
  p.status = o.status;
  p.steps = o.steps;
  p.rejected = o.rejected;
  p.observable = o.observable;
  p.travel_time = o.travel_time;
  p.affine = o.affine;
  p.drift = o.drift;
  p.branch = o.branch;
  p.radiance = o.radiance;
  p.count = o.count;
  return p;
}

}  // namespace

struct gl_trace_plan {
  std::vector<SceneDev> scenes;
  std::vector<SceneHost> host;
  SceneDev* d_scenes = nullptr;
  float* d_map = nullptr;
  uint32_t map_w = 0, map_h = 0;
  QueueDev* d_queue = nullptr;
  FrameScene* d_frame = nullptr;
  FrameScene* h_frame = nullptr;  // pinned staging of the frame table
  size_t frame_capacity = 0;
  std::vector<FrameScene> frame_shadow;  // what d_frame holds (or will hold)
  hipEvent_t frame_copied = nullptr;     // completion of the last staging copy
  bool frame_copy_pending = false;
  DeferRecord* d_defer = nullptr;
  unsigned int* d_defer_count = nullptr;
  uint32_t defer_capacity = 0;
  int cu_count = 256;
  gl_persistent_config pcfg = {256, 2, 8, 16, 2};
  bool timing = false;
  bool capturing = false;
  // Timed launches: a ring of event quadruples (main start/stop, finisher
  // start/stop); a slot is harvested into the sums before it is reused and
  // when the durations are read, so every launch since the last read counts.
  struct TimedLaunch {
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool used = false;
    bool finisher = false;
  };
  static constexpr int kTimedRing = 64;
  TimedLaunch timed[kTimedRing];
  int timed_next = 0;
  int timed_cur = -1;
  double main_sum = 0.0, finish_sum = 0.0;
  uint64_t main_n = 0, finish_n = 0;
};

struct gl_frame_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  FrameScene* d_table = nullptr;
};

namespace {

gl_status ensure_frame(gl_trace_plan* p, size_t n) {
  if (n <= p->frame_capacity) return GL_OK;
  if (p->frame_copy_pending) GL_HIP(hipEventSynchronize(p->frame_copied));
  p->frame_copy_pending = false;
  if (p->d_frame) (void)hipFree(p->d_frame);
  if (p->h_frame) (void)hipHostFree(p->h_frame);
  p->d_frame = nullptr;
  p->h_frame = nullptr;
  p->frame_shadow.clear();
  GL_HIP(hipMalloc(&p->d_frame, n * sizeof(FrameScene)));
  GL_HIP(hipHostMalloc(&p->h_frame, n * sizeof(FrameScene), hipHostMallocDefault));
  if (!p->frame_copied) GL_HIP(hipEventCreateWithFlags(&p->frame_copied, hipEventDisableTiming));
  p->frame_capacity = n;
  return GL_OK;
}

// Stream-ordered upload of the frame table. The device copy is skipped when
// it already holds this table; otherwise the pinned staging buffer is reused
// only after the previous staging copy has executed. Calls on one plan are
// ordered on one stream.
gl_status upload_frame(gl_trace_plan* p, const std::vector<FrameScene>& table, hipStream_t st) {
  gl_status s = ensure_frame(p, table.size());
  if (s) return s;
  if (p->frame_shadow.size() == table.size() &&
      std::memcmp(p->frame_shadow.data(), table.data(), table.size() * sizeof(FrameScene)) == 0)
    return GL_OK;
  if (p->frame_copy_pending) GL_HIP(hipEventSynchronize(p->frame_copied));
  std::memcpy(p->h_frame, table.data(), table.size() * sizeof(FrameScene));
  GL_HIP(hipMemcpyAsync(p->d_frame, p->h_frame, table.size() * sizeof(FrameScene),
                        hipMemcpyHostToDevice, st));
  if (!p->capturing) {
    GL_HIP(hipEventRecord(p->frame_copied, st));
    p->frame_copy_pending = true;
  } else {
    // A captured copy replays from the staging buffer: keep it unchanged.
    p->frame_copy_pending = false;
  }
  p->frame_shadow = table;
  return GL_OK;
}

FrameScene make_frame(uint32_t scene, uint32_t first, uint32_t rays, const gl_ray_outputs& o) {
  FrameScene f;
  std::memset(&f, 0, sizeof f);
  f.scene = scene;
  f.first_item = first;
  f.rays = rays;
  f.out = to_planes(o);
  return f;
}

gl_status ensure_defer(gl_trace_plan* p, size_t rays) {
  // Near-horizon deferral holds at most one record per ray, capped at 4M rays
  // (~1 GB); beyond the cap a ray keeps stepping in f32.
  size_t want = std::min<size_t>(rays, size_t(1) << 22);
  if (want <= p->defer_capacity && p->d_defer) return GL_OK;
  if (p->d_defer) (void)hipFree(p->d_defer);
  p->d_defer = nullptr;
  GL_HIP(hipMalloc(&p->d_defer, want * sizeof(DeferRecord)));
  p->defer_capacity = uint32_t(want);
  return GL_OK;
}

gl_status harvest(gl_trace_plan* p, gl_trace_plan::TimedLaunch& t) {
  if (!t.used) return GL_OK;
  t.used = false;
  float ms = 0.0f;
  GL_HIP(hipEventSynchronize(t.ev[1]));
  GL_HIP(hipEventElapsedTime(&ms, t.ev[0], t.ev[1]));
  p->main_sum += ms;
  ++p->main_n;
  if (t.finisher) {
    GL_HIP(hipEventSynchronize(t.ev[3]));
    GL_HIP(hipEventElapsedTime(&ms, t.ev[2], t.ev[3]));
    p->finish_sum += ms;
    ++p->finish_n;
  }
  return GL_OK;
}

// Event 0 opens a new timed launch (main kernel start), 1 closes the main
// kernel, 2 and 3 bracket the f64 finisher of the same launch.
void mark(gl_trace_plan* p, int which, hipStream_t st) {
  if (!p->timing || p->capturing) return;
  if (which == 0) {
    gl_trace_plan::TimedLaunch& t = p->timed[p->timed_next];
    if (harvest(p, t) != GL_OK) {
      p->timed_cur = -1;
      return;
    }
    p->timed_cur = p->timed_next;
    p->timed_next = (p->timed_next + 1) % gl_trace_plan::kTimedRing;
    t.finisher = false;
  }
  if (p->timed_cur < 0) return;
  gl_trace_plan::TimedLaunch& t = p->timed[p->timed_cur];
  if (hipEventRecord(t.ev[which], st) != hipSuccess) {
    p->timed_cur = -1;
    return;
  }
  if (which == 1) t.used = true;
  if (which == 3) t.finisher = true;
}

MapView map_view(const gl_trace_plan* p) {
  MapView m;
  m.rgb = p->d_map;
  m.w = p->map_w;
  m.h = p->map_h;
  return m;
}

DeferBuf defer_buf(const gl_trace_plan* p) {
  DeferBuf d;
  d.rec = p->d_defer;
  d.count = p->d_defer_count;
  d.capacity = p->defer_capacity;
  return d;
}

bool any_f64_horizon(const gl_trace_plan* p) {
  for (const auto& s : p->scenes)
    if (s.options & kOptF64Horizon) return true;
  return false;
}

gl_status check_outputs(const gl_ray_outputs* o, size_t rays) {
  if (!o) return fail(GL_ERR_INVALID_ARGUMENT, "outputs must not be NULL");
  if (o->count < rays) return fail(GL_ERR_INVALID_ARGUMENT, "output planes hold fewer rays than the scene");
  return GL_OK;
}

// Launch the finisher over however many rays were deferred (read on device).
gl_status launch_finisher(gl_trace_plan* p, hipStream_t st) {
  if (!p->d_defer) return GL_OK;
  mark(p, 2, st);
  hipLaunchKernelGGL(k_finish_f64, dim3(std::max(1, p->cu_count * 4)), dim3(64), 0, st,
                     p->d_scenes, p->d_frame, map_view(p), defer_buf(p));
  GL_HIP(hipGetLastError());
  mark(p, 3, st);
  return GL_OK;
}

// The frame table of a persistent launch: one FrameScene per traced scene
// with the prefix sum of its 8x8-tile work items.
gl_status build_frame_table(gl_trace_plan* p, const uint32_t* scenes, size_t n_scenes,
                            const gl_ray_outputs* outputs, std::vector<FrameScene>* table,
                            unsigned long long* total_items, size_t* total_rays) {
  std::vector<uint32_t> ids;
  if (scenes) {
    ids.assign(scenes, scenes + n_scenes);
  } else {
    for (size_t i = 0; i < p->scenes.size(); ++i) ids.push_back(uint32_t(i));
  }
  if (ids.empty()) return fail(GL_ERR_INVALID_ARGUMENT, "no scenes to trace");
  if (!outputs) return fail(GL_ERR_INVALID_ARGUMENT, "outputs must not be NULL");
  table->resize(ids.size());
  unsigned long long total = 0;
  size_t rays_sum = 0;
  for (size_t i = 0; i < ids.size(); ++i) {
    if (ids[i] >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
    const SceneDev& sd = p->scenes[ids[i]];
    size_t rays = size_t(sd.width) * sd.height;
    gl_status s = check_outputs(&outputs[i], rays);
    if (s) return s;
    uint32_t tiles_y = (sd.height + 7) / 8;
    unsigned long long items = 64ull * sd.tiles_x * tiles_y;
    if (total + items > 0xffffffffull) return fail(GL_ERR_CAPACITY, "frame exceeds 2^32 work items");
    (*table)[i] = make_frame(ids[i], uint32_t(total), uint32_t(rays), outputs[i]);
    total += items;
    rays_sum += rays;
  }
  *total_items = total;
  *total_rays = rays_sum;
  return GL_OK;
}

// Lane-refill kernels for waves_per_simd 2..4; whole-tile kernels (refill
// threshold 64) for 2 and 3.
const void* persistent_kernel(bool single, uint32_t waves, bool tiles) {
  if (tiles) {
    switch (waves) {
      case 2:
        return single ? (const void*)k_persistent_tiles<true, 2> : (const void*)k_persistent_tiles<false, 2>;
      case 3:
        return single ? (const void*)k_persistent_tiles<true, 3> : (const void*)k_persistent_tiles<false, 3>;
      default:
        return nullptr;
    }
  }
  switch (waves) {
    case 2:
      return single ? (const void*)k_persistent<true, 2> : (const void*)k_persistent<false, 2>;
    case 3:
      return single ? (const void*)k_persistent<true, 3> : (const void*)k_persistent<false, 3>;
    case 4:
      return single ? (const void*)k_persistent<true, 4> : (const void*)k_persistent<false, 4>;
    default:
      return nullptr;
  }
}

// Queue reset, persistent kernel and (when any scene asks for f64 horizon
// steps) the f64 finisher, all on `st`, reading the frame table d_table.
gl_status launch_persistent(gl_trace_plan* p, const FrameScene* d_table, size_t n_table,
                            unsigned long long total, size_t total_rays, hipStream_t st) {
  gl_status s = GL_OK;
  bool f64 = any_f64_horizon(p);
  if (f64) {
    s = ensure_defer(p, total_rays);
    if (s) return s;
  }
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, p->d_queue, total,
                     p->d_defer_count);
  GL_HIP(hipGetLastError());
  const gl_persistent_config& c = p->pcfg;
  // Every workgroup must be resident at once: clamp the requested blocks
  // per CU to what the kernel's register budget allows.
  int resident = 0;
  const bool tiles = c.refill_threshold >= 64u;
  const void* kern = persistent_kernel(n_table == 1, c.waves_per_simd, tiles);
  if (!kern)
    return fail(GL_ERR_INVALID_ARGUMENT,
                tiles ? "whole-tile mode (refill_threshold 64) takes waves_per_simd 2 or 3"
                      : "waves_per_simd must be 2, 3 or 4");
  GL_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern, int(c.block_size), 0));
  int per_cu = std::max(1, std::min(int(c.blocks_per_cu), resident));
  dim3 grid(std::max(1, p->cu_count * per_cu));
  dim3 block(c.block_size);
  DeferBuf df = defer_buf(p);
  if (!f64) df.rec = nullptr;
  mark(p, 0, st);
  {
    const SceneDev* a0 = p->d_scenes;
    const FrameScene* a1 = d_table;
    uint32_t a2 = uint32_t(n_table);
    QueueDev* a3 = p->d_queue;
    MapView a4 = map_view(p);
    DeferBuf a5 = df;
    uint32_t a6 = c.refill_interval, a7 = c.refill_threshold;
    void* args[] = {&a0, &a1, &a2, &a3, &a4, &a5, &a6, &a7};
    GL_HIP(hipLaunchKernel(kern, grid, block, args, 0, st));
  }
  GL_HIP(hipGetLastError());
  mark(p, 1, st);
  if (df.rec) {
    mark(p, 2, st);
    hipLaunchKernelGGL(k_finish_f64, dim3(std::max(1, p->cu_count * 4)), dim3(64), 0, st,
                       p->d_scenes, d_table, map_view(p), df);
    GL_HIP(hipGetLastError());
    mark(p, 3, st);
  }
  return GL_OK;
}

gl_status persistent_frame(gl_trace_plan* p, const uint32_t* scenes, size_t n_scenes,
                           const gl_ray_outputs* outputs, hipStream_t st) {
  std::vector<FrameScene> table;
  unsigned long long total = 0;
  size_t rays = 0;
  gl_status s = build_frame_table(p, scenes, n_scenes, outputs, &table, &total, &rays);
  if (s) return s;
  s = upload_frame(p, table, st);
  if (s) return s;
  return launch_persistent(p, p->d_frame, table.size(), total, rays, st);
}

}  // namespace

extern "C" {

const char* gl_last_error(void) { return g_last_error.c_str(); }

const char* gl_version(void) { return "gravlume-mi355x 0.1 (gfx950, f32 DP5(4) trace)"; }

gl_status gl_trace_plan_create(const gl_scene_desc* scenes, size_t n_scenes,
                               const float* escape_map_rgb, uint32_t map_width,
                               uint32_t map_height, gl_trace_plan** out_plan) {
  if (!scenes || !out_plan || n_scenes == 0) return fail(GL_ERR_INVALID_ARGUMENT, "scenes and out_plan are required");
  if (n_scenes > size_t(kMaxScenes)) return fail(GL_ERR_CAPACITY, "more than 4096 scenes");
  *out_plan = nullptr;
  auto* p = new gl_trace_plan();
  p->scenes.resize(n_scenes);
  p->host.resize(n_scenes);
  for (size_t i = 0; i < n_scenes; ++i) {
    std::string err;
    int st = prepare_scene(scenes[i], &p->scenes[i], &p->host[i], &err);
    if (st != GL_OK) {
      delete p;
      return fail(gl_status(st), "scene " + std::to_string(i) + ": " + err);
    }
    if ((scenes[i].options & GL_OPT_ESCAPE_MAP) && !escape_map_rgb) {
      delete p;
      return fail(GL_ERR_INVALID_ARGUMENT, "escape-map option set without a map");
    }
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&p->cu_count, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) {
    delete p;
    return hip_fail(e, "device query");
  }
  auto cleanup = [p](hipError_t er, const char* what) {
    gl_trace_plan_destroy(p);
    return hip_fail(er, what);
  };
  if ((e = hipMalloc(&p->d_scenes, n_scenes * sizeof(SceneDev))) != hipSuccess) return cleanup(e, "hipMalloc scenes");
  if ((e = hipMemcpy(p->d_scenes, p->scenes.data(), n_scenes * sizeof(SceneDev), hipMemcpyHostToDevice)) != hipSuccess)
    return cleanup(e, "upload scenes");
  if ((e = hipMalloc(&p->d_queue, sizeof(QueueDev))) != hipSuccess) return cleanup(e, "hipMalloc queue");
  if ((e = hipMemset(p->d_queue, 0, sizeof(QueueDev))) != hipSuccess) return cleanup(e, "clear queue");
  if ((e = hipMalloc(&p->d_defer_count, 64)) != hipSuccess) return cleanup(e, "hipMalloc defer count");
  if ((e = hipMemset(p->d_defer_count, 0, 64)) != hipSuccess) return cleanup(e, "clear defer count");
  if (escape_map_rgb) {
    if (map_width == 0 || map_height == 0) {
      gl_trace_plan_destroy(p);
      return fail(GL_ERR_INVALID_ARGUMENT, "escape map must be nonempty");
    }
    size_t bytes = size_t(map_width) * map_height * 3 * sizeof(float);
    if ((e = hipMalloc(&p->d_map, bytes)) != hipSuccess) return cleanup(e, "hipMalloc map");
    if ((e = hipMemcpy(p->d_map, escape_map_rgb, bytes, hipMemcpyHostToDevice)) != hipSuccess)
      return cleanup(e, "upload map");
    p->map_w = map_width;
    p->map_h = map_height;
  }
  *out_plan = p;
  return GL_OK;
}

gl_status gl_trace_plan_destroy(gl_trace_plan* p) {
  if (!p) return GL_OK;
  if (p->d_scenes) (void)hipFree(p->d_scenes);
  if (p->d_map) (void)hipFree(p->d_map);
  if (p->d_queue) (void)hipFree(p->d_queue);
  if (p->d_frame) (void)hipFree(p->d_frame);
  if (p->h_frame) (void)hipHostFree(p->h_frame);
  if (p->d_defer) (void)hipFree(p->d_defer);
  if (p->d_defer_count) (void)hipFree(p->d_defer_count);
  if (p->frame_copied) (void)hipEventDestroy(p->frame_copied);
  for (auto& t : p->timed)
    for (hipEvent_t e : t.ev)
      if (e) (void)hipEventDestroy(e);
  delete p;
  return GL_OK;
}

gl_status gl_trace_plan_scene_rays(const gl_trace_plan* p, uint32_t scene, size_t* out_rays) {
  if (!p || !out_rays) return fail(GL_ERR_INVALID_ARGUMENT, "plan and out_rays are required");
  if (scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
  *out_rays = size_t(p->scenes[scene].width) * p->scenes[scene].height;
  return GL_OK;
}

gl_status gl_trace_encode(gl_trace_plan* p, const gl_tile_region* region,
                          const gl_ray_outputs* out, void* stream) {
  if (!p || !region || !out) return fail(GL_ERR_INVALID_ARGUMENT, "plan, region and outputs are required");
  if (region->scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
  const SceneDev& sd = p->scenes[region->scene];
  uint32_t tx_all = sd.tiles_x, ty_all = (sd.height + 7) / 8;
  if (region->tiles_x == 0 || region->tiles_y == 0 || region->tile_x >= tx_all ||
      region->tile_y >= ty_all || region->tiles_x > tx_all - region->tile_x ||
      region->tiles_y > ty_all - region->tile_y)
    return fail(GL_ERR_INVALID_ARGUMENT, "tile region outside the viewport");
  gl_status s = check_outputs(out, size_t(sd.width) * sd.height);
  if (s) return s;
  hipStream_t st = (hipStream_t)stream;
  bool f64 = (sd.options & kOptF64Horizon) != 0;
  if (f64) {
    s = ensure_defer(p, size_t(region->tiles_x) * region->tiles_y * 64);
    if (s) return s;
  }
  s = upload_frame(p, {make_frame(region->scene, 0, sd.width * sd.height, *out)}, st);
  if (s) return s;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, p->d_queue, 0ull, p->d_defer_count);
  DeferBuf df = defer_buf(p);
  if (!f64) df.rec = nullptr;
  size_t blocks = size_t(region->tiles_x) * region->tiles_y;
  mark(p, 0, st);
  hipLaunchKernelGGL(k_plain_image, dim3(uint32_t(blocks)), dim3(64), 0, st, p->d_scenes,
                     p->d_frame, region->tile_x, region->tile_y, region->tiles_x, map_view(p), df);
  GL_HIP(hipGetLastError());
  mark(p, 1, st);
  if (f64) return launch_finisher(p, st);
  return GL_OK;
}

gl_status gl_trace_encode_persistent(gl_trace_plan* p, const uint32_t* scenes, size_t n_scenes,
                                     const gl_ray_outputs* outputs, void* stream) {
  if (!p) return fail(GL_ERR_INVALID_ARGUMENT, "plan is required");
  return persistent_frame(p, scenes, n_scenes, outputs, (hipStream_t)stream);
}

gl_status gl_trace_plan_set_persistent_config(gl_trace_plan* p, const gl_persistent_config* c) {
  if (!p || !c) return fail(GL_ERR_INVALID_ARGUMENT, "plan and config are required");
  if (c->block_size < 64 || c->block_size > 256 || c->block_size % 64 != 0)
    return fail(GL_ERR_INVALID_ARGUMENT, "block_size must be 64, 128, 192 or 256");
  if (c->blocks_per_cu == 0 || c->blocks_per_cu > 32)
    return fail(GL_ERR_INVALID_ARGUMENT, "blocks_per_cu must lie in [1, 32]");
  if (c->refill_interval == 0 || c->refill_interval > 4096)
    return fail(GL_ERR_INVALID_ARGUMENT, "refill_interval must lie in [1, 4096]");
  if (c->refill_threshold == 0 || c->refill_threshold > 64)
    return fail(GL_ERR_INVALID_ARGUMENT, "refill_threshold must lie in [1, 64]");
  if (c->waves_per_simd < 2 || c->waves_per_simd > 4)
    return fail(GL_ERR_INVALID_ARGUMENT, "waves_per_simd must be 2, 3 or 4");
  p->pcfg = *c;
  return GL_OK;
}

gl_status gl_trace_plan_get_persistent_config(const gl_trace_plan* p, gl_persistent_config* c) {
  if (!p || !c) return fail(GL_ERR_INVALID_ARGUMENT, "plan and config are required");
  *c = p->pcfg;
  return GL_OK;
}

gl_status gl_geodesic_batch_trace(gl_trace_plan* p, uint32_t scene, const double* states8,
                                  const float* directions, size_t n, const gl_ray_outputs* out,
                                  void* stream) {
  if (!p || !out) return fail(GL_ERR_INVALID_ARGUMENT, "plan and outputs are required");
  if (scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
  if (n == 0) return GL_OK;  // an empty batch is a no-op (batch.rs:31-43)
  if (!states8 || !directions) return fail(GL_ERR_INVALID_ARGUMENT, "states and directions are required");
  if (n > 0xffffffffull) return fail(GL_ERR_CAPACITY, "more than 2^32 requests");
  gl_status s = check_outputs(out, n);
  if (s) return s;
  hipStream_t st = (hipStream_t)stream;
  bool f64 = (p->scenes[scene].options & kOptF64Horizon) != 0;
  if (f64) {
    s = ensure_defer(p, n);
    if (s) return s;
  }
  s = upload_frame(p, {make_frame(scene, 0, uint32_t(n), *out)}, st);
  if (s) return s;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, p->d_queue, 0ull, p->d_defer_count);
  DeferBuf df = defer_buf(p);
  if (!f64) df.rec = nullptr;
  mark(p, 0, st);
  hipLaunchKernelGGL(k_geodesic_batch, dim3(uint32_t((n + 63) / 64)), dim3(64), 0, st,
                     p->d_scenes, p->d_frame, states8, directions, uint32_t(n), map_view(p), df);
  GL_HIP(hipGetLastError());
  mark(p, 1, st);
  if (f64) return launch_finisher(p, st);
  return GL_OK;
}

gl_status gl_trace_plan_queue_stats(gl_trace_plan* p, gl_queue_stats* stats, void* stream) {
  if (!p || !stats) return fail(GL_ERR_INVALID_ARGUMENT, "plan and stats are required");
  QueueDev q;
  GL_HIP(hipMemcpyAsync(&q, p->d_queue, sizeof(QueueDev), hipMemcpyDeviceToHost, (hipStream_t)stream));
  GL_HIP(hipStreamSynchronize((hipStream_t)stream));
  stats->items = std::min(q.next, q.total);
  stats->refills = q.refills;
  stats->attempted_steps = q.attempts;
  stats->lane_slots = q.slots;
  return GL_OK;
}

gl_status gl_trace_plan_enable_timing(gl_trace_plan* p, int enable) {
  if (!p) return fail(GL_ERR_INVALID_ARGUMENT, "plan is required");
  if (enable) {
    for (auto& t : p->timed)
      for (hipEvent_t& e : t.ev)
        if (!e) GL_HIP(hipEventCreate(&e));
  }
  for (auto& t : p->timed) t.used = false;
  p->timed_cur = -1;
  p->main_sum = p->finish_sum = 0.0;
  p->main_n = p->finish_n = 0;
  p->timing = enable != 0;
  return GL_OK;
}

gl_status gl_trace_plan_kernel_ms(gl_trace_plan* p, double* main_ms, double* finish_ms,
                                  uint64_t* launches) {
  if (!p || !main_ms || !finish_ms) return fail(GL_ERR_INVALID_ARGUMENT, "plan and outputs are required");
  if (!p->timing) return fail(GL_ERR_INVALID_ARGUMENT, "timing is not enabled");
  for (int i = 0; i < gl_trace_plan::kTimedRing; ++i) {
    gl_status s = harvest(p, p->timed[(p->timed_next + i) % gl_trace_plan::kTimedRing]);
    if (s) return s;
  }
  *main_ms = p->main_n ? p->main_sum / double(p->main_n) : 0.0;
  *finish_ms = p->finish_n ? p->finish_sum / double(p->finish_n) : 0.0;
  if (launches) *launches = p->main_n;
  p->main_sum = p->finish_sum = 0.0;
  p->main_n = p->finish_n = 0;
  return GL_OK;
}

gl_status gl_frame_graph_create(gl_trace_plan* p, const uint32_t* scenes, size_t n_scenes,
                                const gl_ray_outputs* outputs, void* stream,
                                gl_frame_graph** out_graph) {
  if (!p || !out_graph) return fail(GL_ERR_INVALID_ARGUMENT, "plan and out_graph are required");
  *out_graph = nullptr;
  hipStream_t st = (hipStream_t)stream;
  if (st == nullptr) return fail(GL_ERR_INVALID_ARGUMENT, "graph capture needs a non-default stream");
  std::vector<FrameScene> table;
  unsigned long long total = 0;
  size_t rays = 0;
  gl_status s = build_frame_table(p, scenes, n_scenes, outputs, &table, &total, &rays);
  if (s) return s;
  if (any_f64_horizon(p)) {
    s = ensure_defer(p, rays);  // allocate before capture
    if (s) return s;
  }
  auto* g = new gl_frame_graph();
  // The graph owns its frame table, written once before capture.
  hipError_t e = hipMalloc(&g->d_table, table.size() * sizeof(FrameScene));
  if (e == hipSuccess)
    e = hipMemcpy(g->d_table, table.data(), table.size() * sizeof(FrameScene), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    gl_frame_graph_destroy(g);
    return hip_fail(e, "graph frame table");
  }
  e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    gl_frame_graph_destroy(g);
    return hip_fail(e, "hipStreamBeginCapture");
  }
  p->capturing = true;
  s = launch_persistent(p, g->d_table, table.size(), total, rays, st);
  p->capturing = false;
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(st, &graph);
  if (s != GL_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    gl_frame_graph_destroy(g);
    return s;
  }
  if (e != hipSuccess) {
    gl_frame_graph_destroy(g);
    return hip_fail(e, "hipStreamEndCapture");
  }
  g->graph = graph;
  e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    gl_frame_graph_destroy(g);
    return hip_fail(e, "hipGraphInstantiate");
  }
  *out_graph = g;
  return GL_OK;
}

gl_status gl_frame_graph_launch(gl_frame_graph* g, void* stream) {
  if (!g || !g->exec) return fail(GL_ERR_INVALID_ARGUMENT, "graph is required");
  GL_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return GL_OK;
}

gl_status gl_frame_graph_destroy(gl_frame_graph* g) {
  if (!g) return GL_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->d_table) (void)hipFree(g->d_table);
  delete g;
  return GL_OK;
}

}  // extern "C"
