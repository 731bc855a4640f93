// Host state of a trace plan, shared by the trace and adjoint entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/gravlume.h"
#include "gl_device.h"
#include "gl_scene.h"
#include "gl_types.h"

namespace gl {
// Records `msg` as the calling thread's last error and returns `st`.
gl_status fail(gl_status st, const std::string& msg);
}  // namespace gl

namespace gl {
// Device buffers of the f64 passes of a launch: the rays deferred to the
// near-horizon finisher with the finisher's claim order, the near-critical
// rays for the refine pass, and their counters (DeferBuf). The plan owns one for its eager calls;
// every frame graph owns its own, so a later call that needs more room can
// reallocate the plan's without touching memory a graph replays into.
struct Workspace {
  DeferRecord* defer = nullptr;
  uint32_t defer_capacity = 0;
  uint32_t* order = nullptr;  // defer_capacity entries
  RefineRecord* refine = nullptr;
  uint32_t refine_capacity = 0;
  uint32_t* refine_order = nullptr;  // refine_capacity entries
  unsigned int* counters = nullptr;  // kCounterWords words
};
}  // namespace gl

#define GL_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::gl::fail(GL_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct gl_trace_plan {
  using SceneDev = gl::SceneDev;
  using SceneHost = gl::SceneHost;
  using QueueDev = gl::QueueDev;
  using FrameScene = gl::FrameScene;
  using DeferRecord = gl::DeferRecord;
  std::vector<SceneDev> scenes;
  std::vector<SceneHost> host;
  SceneDev* d_scenes = nullptr;
  float* d_map = nullptr;
  float* d_lut = nullptr;  // blackbody band table, when a scene needs it
  uint32_t map_w = 0, map_h = 0;
  QueueDev* d_queue = nullptr;
  FrameScene* d_frame = nullptr;
  FrameScene* h_frame = nullptr;  // pinned staging of the frame table
  size_t frame_capacity = 0;
  std::vector<FrameScene> frame_shadow;  // what d_frame holds (or will hold)
  hipEvent_t frame_copied = nullptr;     // completion of the last staging copy
  bool frame_copy_pending = false;
  gl::Workspace ws;             // deferral and refine buffers of eager calls
  uint32_t* d_tiles = nullptr;  // per-scene tile orders, then tile costs
  int cu_count = 256;
  // whole-tile mode: the faster one on every workload measured
  gl_persistent_config pcfg = {256, 2, 8, 64, 2};
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
  // Escape-direction stencil nodes and the shadow-edge pixel list of the
  // accelerated sky trace and coverage refinement (gl_accel.hip).
  float4* d_nodes = nullptr;
  uint32_t* d_ready = nullptr;  // per-node epoch flags of the fused stencil launch
  uint32_t sky_epoch = 0;
  size_t node_capacity = 0;
  uint32_t* d_edges = nullptr;  // [0] count, [1..] pixel indices
  size_t edge_capacity = 0;
  // f64 overlap (gl_trace_plan_set_f64_overlap): the f64 passes of an eager
  // persistent launch run on the plan's second stream `side`, beside the
  // next launch's trace. Launches alternate between two workspaces (ws and
  // ws_alt) and two copies of the frame table the f64 passes read; a
  // launch's passes are joined into its caller's stream after the next
  // launch's trace kernel, or by any other entry point (join_side) or
  // gl_trace_plan_flush.
  bool overlap = false;
  hipStream_t side = nullptr;
  gl::Workspace ws_alt;
  FrameScene* d_frame_ov[2] = {nullptr, nullptr};
  size_t frame_ov_capacity[2] = {0, 0};
  std::vector<FrameScene> frame_ov_shadow[2];
  hipEvent_t ev_main[2] = {nullptr, nullptr};
  hipEvent_t ev_fin[2] = {nullptr, nullptr};
  bool fin_recorded[2] = {false, false};  // ev_fin[k] recorded, not yet waited for by the host
  bool fin_unjoined[2] = {false, false};  // ev_fin[k] not yet waited for by the caller's stream
  int ov_parity = 0;
};


namespace gl {
// Makes `st` wait for every f64 pass still running on the plan's second
// stream (f64 overlap), so that work enqueued on `st` afterwards may reuse
// the plan's workspaces and see those passes' outputs.
gl_status join_side(gl_trace_plan* p, hipStream_t st);
// The host waits until no f64 pass is in flight on the plan's second stream
// (before an entry point reallocates or rewrites the workspaces, captures a
// graph or destroys the plan).
gl_status drain_side(gl_trace_plan* p);

// Records timing event `which` (0/1 main kernel start/stop, 2/3 finisher)
// of the current launch on `st` when the plan's timing is enabled.
void plan_mark(gl_trace_plan* p, int which, hipStream_t st);

// Host helpers of the trace entry points (gl_trace.hip), shared with the
// accelerated sky trace (gl_accel.hip).
gl_status hip_fail(hipError_t e, const char* what);
gl_status check_outputs(const gl_ray_outputs* o, size_t rays);
OutPlanes to_planes(const gl_ray_outputs& o);
FrameScene make_frame(uint32_t scene, uint32_t first, uint32_t rays, const gl_ray_outputs& o);
gl_status upload_frame(gl_trace_plan* p, const std::vector<FrameScene>& table, hipStream_t st);
// Grows `w` to hold the deferral records (f64) and refine records (refine)
// of a launch of `rays` rays, each capped at 4M records.
gl_status ensure_workspace(Workspace& w, size_t rays, bool f64, bool refine);
void free_workspace(Workspace& w);
// The kernels' view of `w`: no deferral unless f64, no refine pass unless refine.
DeferBuf workspace_view(const Workspace& w, bool f64, bool refine);
// Whether any scene of the plan (or scene `scene`) defers near-horizon
// steps to f64, and whether it refines near-critical rays.
bool any_f64_horizon(const gl_trace_plan* p);
bool any_refine(const gl_trace_plan* p);
bool scene_f64_horizon(const SceneDev& s);
bool scene_refine(const SceneDev& s);
// The near-horizon finisher (when df defers) and the refine pass (when df
// refines) after a trace launch on `st`, reading the frame table `table`.
gl_status launch_f64_passes(gl_trace_plan* p, const FrameScene* table, const DeferBuf& df,
                            hipStream_t st);
// The next whole-tile launch's tile order from this launch's tile costs, on
// `st` after the launch.
gl_status launch_tile_order(gl_trace_plan* p, const FrameScene* d_table, size_t n_table,
                            hipStream_t st);
MapView map_view(const gl_trace_plan* p);

__global__ void k_queue_reset(QueueDev* q, unsigned long long total,
                              unsigned long long batch_first, unsigned int* defer_count);
}  // namespace gl
