// Device-resident scene record and launch bookkeeping shared by the host
// packer (gl_scene.cpp) and the kernels (gl_trace.hip).
#pragma once

#include <stdint.h>

namespace gl {

// One scene as the kernels read it: the f32 image of the validated f64 scene
// plus the f64 values the f64 islands, the f64 initial rays, the
// near-critical test and the surface radiance need. 528 bytes, read through
// the scalar cache (the role of TraceUniforms, trace/input.rs:19-35).
struct alignas(16) SceneDev {
  // spacetime
  float mass, spin, spin2, charge2;
  float chart;          // +1 ingoing, -1 outgoing
  float r_capture;      // r_+ (1 + offset); < 0 without a horizon
  float r_horizon;      // r_+; < 0 without a horizon
  float r_escape;
  // surface
  float disk_in, disk_out;  // disk_in < 0 disables
  uint32_t emission_model;  // 0 bolometric, 1 blackbody
  float arming_band;
  // policy
  float rtol, atol;
  float h_init, h_min, h_max, h_max_per_r;
  float tie_tol, drift_limit;
  float guard;  // singularity guard on r^4 + a^2 z^2
  uint32_t max_accepted, max_rejects;
  uint32_t options;
  uint32_t reproject_interval;
  // observer and view
  float obs_x, obs_y, obs_z;
  float u_t;                        // observer four-velocity, time component
  float right[4], up[4], arrival[4];  // tetrad legs (t, x, y, z)
  float tan_half_fov, aspect;
  float sub_x, sub_y;
  float omega;
  uint32_t width, height;
  uint32_t polar_side;  // 0 negative, 1 equatorial, 2 positive
  uint32_t ray_offset;  // first work item of this scene in a persistent frame
  uint32_t tiles_x;     // ceil(width / 8)
  // f64 copies for the f64 islands
  double spin_d, charge2_d, r_horizon_d;
  double r_capture_d, disk_in_d, disk_out_d;
  // surface radiance (trace/input.rs:166-320): emitter and slab transport
  double intensity6_d;     // bolometric intensity at r = 6
  double temperature6_d;   // blackbody temperature at r = 6 [K]
  double transmittance_d;  // exp(-optical depth), 1 in vacuum
  double slab_emission_d;  // integrated slab emission (0 in vacuum)
  double slab_temperature_d;  // its blackbody temperature [K], 0 neutral
  const float* lut;        // 4097 x 4 log2 band fractions (blackbody scenes)
  float safety, min_factor, max_factor;  // step controller factors
  uint32_t pad_;
  // whole-tile persistent queue: the order in which the tiles are handed out
  // (longest of the previous launch first) and this launch's cost per tile
  const uint32_t* tile_order;
  uint32_t* tile_cost;
  // f64 observer event and tetrad: every pixel's initial ray is formed in
  // f64 (the f32 trace starts from its rounding; the f64 refine from it)
  double obs_d[3];
  double u_t_d;
  double right_d[4], up_d[4], arrival_d[4];
  double tan_half_fov_d, aspect_d, sub_x_d, sub_y_d;
  double r_observer_d;  // oblate radius of the observer
  double phi_observer_d;  // azimuth atan2(y, x) of the observer
  // near-critical test and its f64 refine (gl_scene_desc)
  float near_critical_factor;  // <= 0: no test
  float branch_factor;
  float refine_rtol, refine_atol;
};

static_assert(sizeof(SceneDev) == 528, "SceneDev must stay 528 bytes");

// Work-queue and statistics block of one persistent launch.
struct alignas(64) QueueDev {
  unsigned long long next;      // next unclaimed work item
  unsigned long long total;     // items in the frame
  unsigned long long refills;
  unsigned long long attempts;  // DP5 attempts by active lanes
  unsigned long long slots;     // lane-attempt slots offered
  unsigned long long aux;       // second cursor (the accelerated sky trace's nodes)
  unsigned long long order_dirty;  // a whole-tile launch saw a tile cost change
  // Whole-tile mode over a mixed work list: items from batch_first on are
  // geodesic-batch requests, claimed lane by lane (from batch_next) once the
  // image tiles are handed out; batch_first = total when there are none.
  unsigned long long batch_first;
  unsigned long long batch_next;
  unsigned long long pad[7];
};
static_assert(sizeof(QueueDev) == 128, "QueueDev is two 64-byte lines");

// Output planes as the kernels see them (see gl_ray_outputs in gravlume.h).
struct OutPlanes {
  uint32_t* status;
  uint32_t* steps;
  uint32_t* rejected;
  float* observable;
  float* travel_time;
  float* affine;
  float* drift;
  uint32_t* branch;
  float* radiance;
  uint16_t* rgba16f;  // 4 halves per ray
  float* event_residual;
  float* turning_radius;
  float* azimuth_advance;
  unsigned long long count;
};

// Per-scene slice of a persistent frame.
struct FrameScene {
  uint32_t scene;       // index into the scene table
  uint32_t first_item;  // prefix sum of queue items
  uint32_t rays;
  uint32_t pad;
  OutPlanes out;
  // Geodesic batch: f64 initial states (8 per request) and affine directions;
  // null for image scenes, whose items are 8x8-tile-ordered pixels.
  const double* states;
  const float* dirs;
};

constexpr int kMaxScenes = 4096;

}  // namespace gl
