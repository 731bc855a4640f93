// Image-space accelerators around the trace:
//
//   * the escape-direction stencil (the reference's AcceleratedSky plan,
//     crates/gravlume-render/src/shaders/geodesic_acceleration.wgsl:330-601):
//     every 4th pixel in each direction (the nodes) is traced in full first;
//     an 8x8 tile whose 3x3 nodes all escape, and whose edge-midpoint and
//     centre nodes lie within min(half a pixel, 3e-4) of the bilinear blend
//     of its corner nodes, takes the escape direction of its other pixels
//     from that blend instead of tracing them. Every other tile is traced in
//     full. Sky scenes only (no surface).
//   * shadow-edge coverage (shadow_coverage.wgsl:1-136): pixels whose
//     capture/escape class differs from a 4-neighbour's are listed, and each
//     is re-traced at four rotated-grid sub-samples; the texel becomes the
//     mean escape colour with alpha = escaping fraction.
//
// MI355X layout: one launch of persistent 64-lane wavefronts drains a work
// queue whose first items are 8x8 blocks of nodes and whose later items are
// 8x8 tiles, so the long tail of the node traces overlaps the tile work
// instead of ending a separate pass. A tile's wave waits (bounded, with
// s_sleep) for its nine node records, published with agent-scope release
// stores behind a per-launch epoch flag: every node item is claimed before
// any tile item, so the nodes a tile waits for are already being traced by
// resident waves. The stencil is one LDS vector per wave and the accept
// decision is wave-uniform; node pixels keep their exact full-trace outputs,
// so no quantized direction is ever reused. Refinement packs the four
// sub-samples of a pixel into a lane quad and reduces them with shuffles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/gravlume.h"
#include "gl_device.h"
#include "gl_plan.h"
#include "gl_types.h"
#include "trace_device.h"

namespace gl {

constexpr float kNodeOther = 0.0f;
constexpr float kNodeEscape = 1.0f;
constexpr float kNodeHorizon = 2.0f;
constexpr float kStencilMaxError = 3.0e-4f;  // rad; half a pixel caps it further

constexpr uint32_t kMaxNodeWait = 1u << 16;  // s_sleep rounds before a tile stops waiting

__global__ void k_accel_reset(unsigned int* defer_count, uint32_t* edge_count, QueueDev* q) {
  if (blockIdx.x != 0) return;
  if (defer_count) clear_counters(defer_count);
  if (threadIdx.x == 0) {
    if (edge_count) edge_count[0] = 0;
    if (q) {
      q->next = 0;
      q->aux = 0;
    }
  }
}

__device__ __forceinline__ float3 blend(float4 a, float4 b, float t) {
  return make_float3(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t);
}

__device__ __forceinline__ float3 unit(float3 v) {
  const float inv = 1.0f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return make_float3(v.x * inv, v.y * inv, v.z * inv);
}

// Bilinear blend of the four corner nodes at tile position (u, v) in [0, 1].
__device__ __forceinline__ float3 corner_blend(const float4* st, float u, float v) {
  const float3 up = blend(st[0], st[2], u), lo = blend(st[6], st[8], u);
  return unit(make_float3(up.x + (lo.x - up.x) * v, up.y + (lo.y - up.y) * v,
                          up.z + (lo.z - up.z) * v));
}

__device__ __forceinline__ bool node_matches(const float4* st, int k, float u, float v,
                                             float max_err2) {
  const float3 p = corner_blend(st, u, v);
  const float ex = p.x - st[k].x, ey = p.y - st[k].y, ez = p.z - st[k].z;
  const float e2 = ex * ex + ey * ey + ez * ez;
  return e2 <= max_err2;  // false for a non-finite blend
}

// Outputs of a pixel whose escape direction comes from the stencil: class
// escape with candidate bit 8 and the reconstructed flag; the traced-only
// quantities (steps, travel time, affine parameter, drifts, turnings,
// crossings, winding, event residual, azimuth advance) are zero and the
// turning radius -1.
__device__ void write_reconstructed(const OutPlanes& o, uint32_t idx, const SceneDev& s,
                                    const MapView& map, float3 d) {
  const unsigned long long n = o.count;
  if (o.status) o.status[idx] = kTermEscape | (kEvEsc << 4) | (kFlagReconstructed << 8);
  if (o.steps) o.steps[idx] = 0u;
  if (o.rejected) o.rejected[idx] = 0u;
  if (o.observable) {
    o.observable[idx] = d.x;
    o.observable[n + idx] = d.y;
    o.observable[2 * n + idx] = d.z;
  }
  if (o.travel_time) o.travel_time[idx] = 0.0f;
  if (o.affine) o.affine[idx] = 0.0f;
  if (o.drift) {
    o.drift[idx] = 0.0f;
    o.drift[n + idx] = 0.0f;
    o.drift[2 * n + idx] = 0.0f;
  }
  if (o.branch) o.branch[idx] = (s.polar_side & 3u) << 24;
  if (o.event_residual) o.event_residual[idx] = 0.0f;
  if (o.turning_radius) o.turning_radius[idx] = -1.0f;
  if (o.azimuth_advance) o.azimuth_advance[idx] = 0.0f;
  if (o.radiance || o.rgba16f) {
    const Rgb c = escape_texel(map, s, d.x, d.y, d.z);
    if (o.radiance) {
      o.radiance[idx] = c.r;
      o.radiance[n + idx] = c.g;
      o.radiance[2 * n + idx] = c.b;
    }
    if (o.rgba16f) store_rgba16f(o.rgba16f, idx, c, 1.0f);
  }
}

// Publishes the stencil record of one node (after its pixel's outputs).
__device__ __forceinline__ void publish_node(float4* nodes, uint32_t* ready, uint32_t node,
                                             float4 rec, uint32_t epoch) {
  nodes[node] = rec;
  __hip_atomic_store(ready + node, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// The stencil record of a finished node ray: class tag and unit escape
// direction, read back from the observable plane its outputs went to.
__device__ __forceinline__ float4 node_record(const RayT<float>& R, const OutPlanes& o,
                                              uint32_t idx) {
  if (R.term == kTermHorizon) return make_float4(0.0f, 0.0f, 0.0f, kNodeHorizon);
  if (R.term == kTermEscape && !(R.flags & kFlagTied)) {
    const unsigned long long n = o.count;
    const float dx = o.observable[idx], dy = o.observable[n + idx], dz = o.observable[2 * n + idx];
    if (dx * dx + dy * dy + dz * dz > 0.5f) return make_float4(dx, dy, dz, kNodeEscape);
  }
  return make_float4(0.0f, 0.0f, 0.0f, kNodeOther);
}

// The fused stencil launch: persistent wavefronts take queue items; items
// below node_items are 8x8 node blocks (pixel (4 nx, 4 ny) per lane, traced
// in full, its record published to the tiles), the rest 8x8 tiles. Both
// kinds share one trace loop whose lanes all start together, so a traced
// pixel comes out exactly as in the plain dispatch. Every wave reaches the
// end of the queue.
__global__ void __launch_bounds__(64, 2) k_sky_fused(
    const SceneDev* __restrict__ scenes, const FrameScene* __restrict__ frame, QueueDev* q,
    uint32_t node_w, uint32_t node_h, uint32_t node_blocks_x, uint32_t node_items,
    uint32_t tiles_x, uint32_t tiles, float4* __restrict__ nodes, uint32_t* __restrict__ ready,
    uint32_t epoch, float max_err2, MapView map, DeferBuf df) {
  __shared__ float4 st[9];
  const FrameScene& fsc = frame[0];
  const SceneDev& s = scenes[fsc.scene];
  SceneT<float> sc = scene_f32(s);
  hold_in_vgprs(sc);
  const uint32_t lane = threadIdx.x;
  const uint32_t lx = lane & 7u, ly = lane >> 3;
  const uint32_t total = node_items + tiles;
  const float4 other = make_float4(0.0f, 0.0f, 0.0f, kNodeOther);
  for (;;) {
    uint32_t item = 0u;
    if (lane == 0u) item = atomicAdd(reinterpret_cast<unsigned int*>(&q->next), 1u);
    item = __shfl(item, 0);
    if (item >= total) break;
    const bool node_item = item < node_items;
    uint32_t px = 0u, py = 0u, node = 0u;
    bool node_lane = false, trace = false;
    if (node_item) {
      const uint32_t nx = (item % node_blocks_x) * 8u + lx;
      const uint32_t ny = (item / node_blocks_x) * 8u + ly;
      node_lane = nx < node_w && ny < node_h;
      node = ny * node_w + nx;
      px = nx * 4u;
      py = ny * 4u;
      trace = node_lane && px < s.width && py < s.height;
    } else {
      const uint32_t tile = item - node_items;
      const uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
      const uint32_t nidx = lane < 9u ? (2u * ty + lane / 3u) * node_w + 2u * tx + lane % 3u : 0u;
      bool have = false;
      for (uint32_t w = 0; w < kMaxNodeWait; ++w) {
        const bool ok = lane >= 9u || __hip_atomic_load(ready + nidx, __ATOMIC_ACQUIRE,
                                                         __HIP_MEMORY_SCOPE_AGENT) == epoch;
        if (__all(ok)) {
          have = true;
          break;
        }
        __builtin_amdgcn_s_sleep(4);
      }
      if (lane < 9u) st[lane] = have ? nodes[nidx] : other;
      __syncthreads();
      bool accept = true;
#pragma unroll
      for (int k = 0; k < 9; ++k) accept = accept && st[k].w == kNodeEscape;
      accept = accept && node_matches(st, 1, 0.5f, 0.0f, max_err2) &&
               node_matches(st, 3, 0.0f, 0.5f, max_err2) &&
               node_matches(st, 4, 0.5f, 0.5f, max_err2) &&
               node_matches(st, 5, 1.0f, 0.5f, max_err2) && node_matches(st, 7, 0.5f, 1.0f, max_err2);
      px = tx * 8u + lx;
      py = ty * 8u + ly;
      const bool node_pixel = (lx & 3u) == 0u && (ly & 3u) == 0u;  // already final
      const bool inside = px < s.width && py < s.height && !node_pixel;
      if (inside && accept) {
        const uint32_t cx = min(lx >> 2, 1u), cy = min(ly >> 2, 1u);
        const float fu = float(lx - 4u * cx) * 0.25f, fv = float(ly - 4u * cy) * 0.25f;
        const uint32_t first = cy * 3u + cx;
        const float3 up = blend(st[first], st[first + 1], fu);
        const float3 lo = blend(st[first + 3], st[first + 4], fu);
        const float3 d = unit(make_float3(up.x + (lo.x - up.x) * fv, up.y + (lo.y - up.y) * fv,
                                          up.z + (lo.z - up.z) * fv));
        write_reconstructed(fsc.out, py * s.width + px, s, map, d);
      }
      trace = inside && !accept;
    }
    if (trace) {
      const uint32_t idx = py * s.width + px;
      RayT<float> R;
      init_image_ray(s, sc, px, py, R);
      bool deferred = false;
      uint32_t it = 1;
      trace_f32(sc, R, s.omega, df, 0u, idx, it, deferred);
      // a deferred or refined node is finished in f64 later: not a stencil node
      if (!deferred)
        deferred = !write_ray(R, fsc.out, idx, s, map, pixel_source(px, py), refine_sink(df), 0u);
      if (node_item)
        publish_node(nodes, ready, node, deferred ? other : node_record(R, fsc.out, idx), epoch);
    } else if (node_lane) {
      publish_node(nodes, ready, node, other, epoch);  // outside the image
    }
    __syncthreads();  // st is rewritten by the next tile
  }
}

// Shadow-edge classification: capture (0) / escape (1) against the four
// neighbours; edge pixels are appended to edges[1..capacity].
__device__ __forceinline__ int shadow_class(const uint32_t* status, uint32_t i) {
  const uint32_t t = status[i] & 15u;
  return t == kTermHorizon ? 0 : (t == kTermEscape ? 1 : -1);
}

__global__ void k_shadow_classify(const uint32_t* __restrict__ status, uint32_t w, uint32_t h,
                                  uint32_t* __restrict__ edges, uint32_t capacity) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  const int c = shadow_class(status, i);
  if (c < 0) return;
  const uint32_t x = i % w, y = i / w;
  const int o = 1 - c;
  const bool edge = (x > 0u && shadow_class(status, i - 1u) == o) ||
                    (x + 1u < w && shadow_class(status, i + 1u) == o) ||
                    (y > 0u && shadow_class(status, i - w) == o) ||
                    (y + 1u < h && shadow_class(status, i + w) == o);
  if (!edge) return;
  const uint32_t slot = atomicAdd(edges, 1u);
  if (slot < capacity) edges[1u + slot] = i;
}

// Refinement: a lane quad per edge pixel, one rotated-grid sub-sample per
// lane, f32 to the terminal (no f64 deferral); waves stride over the list.
__global__ void __launch_bounds__(64) k_shadow_refine(const SceneDev* __restrict__ scenes,
                                                      const FrameScene* __restrict__ frame,
                                                      const uint32_t* __restrict__ edges,
                                                      uint32_t capacity, MapView map) {
  const FrameScene& fsc = frame[0];
  const SceneDev& s = scenes[fsc.scene];
  SceneT<float> sc = scene_f32(s);
  hold_in_vgprs(sc);
  // the four-sample rotated grid (x, y) in eighths of a pixel
  const float kOff[4][2] = {{0.375f, 0.125f}, {0.875f, 0.375f}, {0.625f, 0.875f}, {0.125f, 0.625f}};
  const uint32_t count = min(edges[0], capacity);
  const uint32_t q = threadIdx.x >> 2, sub = threadIdx.x & 3u;
  for (uint32_t base = blockIdx.x * 16u; base < count; base += gridDim.x * 16u) {
    const uint32_t e = base + q;
    const bool active = e < count;
    const uint32_t idx = active ? edges[1u + e] : 0u;
    float cr = 0.0f, cg = 0.0f, cb = 0.0f;
    uint32_t escapes = 0u, invalid = active ? 0u : 1u;
    if (active) {
      const uint32_t px = idx % s.width, py = idx / s.width;
      RayT<float> R;
      init_image_ray_at(s, sc, px, py, kOff[sub][0], kOff[sub][1], R);
      for (uint32_t it = 1; R.term == kTermActive; ++it)
        step<float>(sc, R, s.omega, reproject_now(sc, it), drift_now(it));
      if (R.term == kTermEscape) {
        const Dir3 d = escape_direction(s, R.dir, R.x, R.y, R.z, R.px, R.py, R.pz, R.E);
        if (d.ok) {
          const Rgb c = escape_texel(map, s, d.x, d.y, d.z);
          cr = c.r;
          cg = c.g;
          cb = c.b;
          escapes = 1u;
        } else {
          invalid = 1u;
        }
      } else if (R.term != kTermHorizon) {
        invalid = 1u;
      }
    }
    // sum over the quad
#pragma unroll
    for (int m = 1; m <= 2; m <<= 1) {
      cr += __shfl_xor(cr, m);
      cg += __shfl_xor(cg, m);
      cb += __shfl_xor(cb, m);
      escapes += __shfl_xor(escapes, m);
      invalid |= __shfl_xor(invalid, m);
    }
    if (sub == 0u && active && invalid == 0u && fsc.out.rgba16f)
      store_rgba16f(fsc.out.rgba16f, idx, Rgb{0.25f * cr, 0.25f * cg, 0.25f * cb},
                    0.25f * float(escapes));
  }
}

// Source-chart footprint (ObservationTracer::surface_footprint_v1,
// crates/gravlume-reference/src/footprint.rs:74-151, and its GPU capture
// surface_footprint_capture.wgsl:1-99): per requested pixel, a group of
// five lanes (twelve groups per wavefront) traces the centre and the four
// +-0.25-pixel neighbours, the group leader checks that all five end unambiguously on the
// surface (flags 0, candidate mask exactly the surface bit) with one branch
// key, and forms the central-difference Jacobian of (r, r_c dphi) against
// the pixel coordinates, with its parity.
constexpr uint32_t kFootprintGroups = 12u;

__global__ void __launch_bounds__(64) k_surface_footprint(const SceneDev* __restrict__ scenes,
                                                          const FrameScene* __restrict__ frame,
                                                          const uint32_t* __restrict__ pixels,
                                                          uint32_t n, float* __restrict__ jac,
                                                          uint32_t* __restrict__ meta) {
  const FrameScene& fsc = frame[0];
  const SceneDev& s = scenes[fsc.scene];
  SceneT<float> sc = scene_f32(s);
  hold_in_vgprs(sc);
  const float kOff[5][2] = {{0.0f, 0.0f}, {-0.25f, 0.0f}, {0.25f, 0.0f}, {0.0f, -0.25f}, {0.0f, 0.25f}};
  // twelve five-lane groups per wavefront (lanes 60-63 idle): the centre and
  // its four neighbours of one pixel per group, gathered at the group's
  // first lane
  const uint32_t group = threadIdx.x / 5u, k = threadIdx.x % 5u;
  const uint32_t lead = min(group, kFootprintGroups - 1u) * 5u;
  const bool in_group = group < kFootprintGroups;
  for (uint32_t base = blockIdx.x * kFootprintGroups; base < n;
       base += gridDim.x * kFootprintGroups) {
    const uint32_t e = in_group ? base + group : n;
    float r = 0.0f, phi = 0.0f;
    uint32_t key = 0u, ok = 0u, term = 0u, flags = 0u;
    // an index outside the image is not traced: zeros, meta 0
    const uint32_t pix = e < n ? (pixels ? pixels[e] : e) : 0u;
    const bool inside = uint64_t(pix) < uint64_t(s.width) * s.height;
    if (e < n && inside) {
      const uint32_t px = pix % s.width, py = pix / s.width;
      RayT<float> R;
      init_image_ray_at(s, sc, px, py, s.sub_x + kOff[k][0], s.sub_y + kOff[k][1], R);
      for (uint32_t it = 1; R.term == kTermActive; ++it)
        step<float>(sc, R, s.omega, reproject_now(sc, it), drift_now(it));
      term = R.term;
      flags = R.flags;
      key = min(R.turnings, 255u) | (min(R.crossings, 255u) << 8) |
            ((uint32_t(R.winding) & 255u) << 16) | ((R.polar_side & 3u) << 24);
      if (R.term == kTermSurface && R.flags == 0u && R.candidates == kEvSurf) {
        const Anchor a = surface_anchor(s, R.x, R.y, R.z, R.cx, R.cy, R.cz, R.px, R.py, R.E);
        r = a.r;
        phi = a.phi;
        ok = a.ok ? 1u : 0u;
      }
    }
    // gather the five samples at the group's first lane
    float rs[5], ps[5];
    uint32_t keys[5], oks = 0u;
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      rs[j] = __shfl(r, int(lead) + j);
      ps[j] = __shfl(phi, int(lead) + j);
      keys[j] = __shfl(key, int(lead) + j);
      oks |= __shfl(ok, int(lead) + j) << j;
    }
    if (k == 0u && e < n) {
      float J[4] = {0.0f, 0.0f, 0.0f, 0.0f};
      uint32_t resolved = 0u, parity = 0u;
      bool same = true;
#pragma unroll
      for (int j = 1; j < 5; ++j) same = same && keys[j] == keys[0];
      if (oks == 31u && same) {
        const float pi = 3.14159265358979f, tau = 6.28318530717959f;
        float arc[5];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
          const float d = ps[j] - ps[0];
          arc[j] = rs[0] * (d - tau * floorf((d + pi) / tau));
        }
        J[0] = (rs[2] - rs[1]) * 2.0f;
        J[1] = (rs[4] - rs[3]) * 2.0f;
        J[2] = (arc[2] - arc[1]) * 2.0f;
        J[3] = (arc[4] - arc[3]) * 2.0f;
        const float det = J[0] * J[3] - J[1] * J[2];
        const float norm2 = J[0] * J[0] + J[1] * J[1] + J[2] * J[2] + J[3] * J[3];
        const float floor_ = 64.0f * 1.1920929e-7f * norm2;
        parity = det > floor_ ? 1u : (det < -floor_ ? 2u : 3u);
        resolved = (norm2 == norm2 && norm2 <= 3.0e38f) ? 1u : 0u;
        if (!resolved) J[0] = J[1] = J[2] = J[3] = 0.0f;
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) jac[4 * size_t(e) + c] = J[c];
      meta[e] = resolved | (parity << 1) | ((term & 15u) << 8) | ((flags & 255u) << 16);
    }
  }
}

}  // namespace gl

using namespace gl;

namespace {

gl_status scene_target(gl_trace_plan* p, uint32_t scene, const gl_ray_outputs* out,
                       const SceneDev** sd) {
  if (!p) return fail(GL_ERR_INVALID_ARGUMENT, "plan is required");
  if (scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
  *sd = &p->scenes[scene];
  return check_outputs(out, size_t((*sd)->width) * (*sd)->height);
}

}  // namespace

extern "C" {

gl_status gl_trace_encode_accelerated(gl_trace_plan* p, uint32_t scene, const gl_ray_outputs* out,
                                      void* stream) {
  const SceneDev* sd = nullptr;
  gl_status s = scene_target(p, scene, out, &sd);
  if (s) return s;
  if (sd->disk_in > 0.0f)
    return fail(GL_ERR_INVALID_ARGUMENT,
                "the escape-direction stencil applies to scenes without a surface");
  if (!out->observable) return fail(GL_ERR_INVALID_ARGUMENT, "the stencil needs the observable plane");
  hipStream_t st = (hipStream_t)stream;
  s = drain_side(p);  // the workspace below may be reallocated and is rewritten
  if (s) return s;
  const uint32_t tiles_x = (sd->width + 7u) / 8u, tiles_y = (sd->height + 7u) / 8u;
  const uint32_t node_w = 2u * tiles_x + 1u, node_h = 2u * tiles_y + 1u;
  const size_t nodes = size_t(node_w) * node_h;
  if (nodes > p->node_capacity) {
    if (p->d_nodes) (void)hipFree(p->d_nodes);
    if (p->d_ready) (void)hipFree(p->d_ready);
    p->d_nodes = nullptr;
    p->d_ready = nullptr;
    p->node_capacity = 0;
    GL_HIP(hipMalloc(&p->d_nodes, nodes * sizeof(float4)));
    GL_HIP(hipMalloc(&p->d_ready, nodes * sizeof(uint32_t)));
    // no flag may hold a future epoch: start from zero (epochs start at 1)
    GL_HIP(hipMemset(p->d_ready, 0, nodes * sizeof(uint32_t)));
    p->node_capacity = nodes;
    p->sky_epoch = 0;
  }
  if (++p->sky_epoch == 0u) {
    // 2^32 launches: clear the flags again before the epoch restarts
    GL_HIP(hipMemsetAsync(p->d_ready, 0, p->node_capacity * sizeof(uint32_t), st));
    p->sky_epoch = 1u;
  }
  const bool f64 = scene_f64_horizon(*sd), refine = scene_refine(*sd);
  s = ensure_workspace(p->ws, size_t(sd->width) * sd->height, f64, refine);
  if (s) return s;
  s = upload_frame(p, {make_frame(scene, 0, sd->width * sd->height, *out)}, st);
  if (s) return s;
  hipLaunchKernelGGL(k_accel_reset, dim3(1), dim3(64), 0, st, p->ws.counters, (uint32_t*)nullptr,
                     p->d_queue);
  const DeferBuf df = workspace_view(p->ws, f64, refine);
  const double half_pixel = double(p->host[scene].tan_half_fov) / double(sd->height);
  const float max_err = float(std::min(half_pixel, double(kStencilMaxError)));
  const uint32_t blocks_x = (node_w + 7u) / 8u, blocks_y = (node_h + 7u) / 8u;
  const uint32_t node_items = blocks_x * blocks_y, tiles = tiles_x * tiles_y;
  const uint32_t waves = std::min<uint32_t>(uint32_t(std::max(1, p->cu_count)) * 8u, node_items + tiles);
  plan_mark(p, 0, st);
  hipLaunchKernelGGL(k_sky_fused, dim3(waves), dim3(64), 0, st, p->d_scenes, p->d_frame, p->d_queue,
                     node_w, node_h, blocks_x, node_items, tiles_x, tiles, p->d_nodes, p->d_ready,
                     p->sky_epoch, max_err * max_err, map_view(p), df);
  GL_HIP(hipGetLastError());
  plan_mark(p, 1, st);
  return launch_f64_passes(p, p->d_frame, df, st);
}

gl_status gl_trace_refine_shadow(gl_trace_plan* p, uint32_t scene, const gl_ray_outputs* out,
                                 void* stream) {
  const SceneDev* sd = nullptr;
  gl_status s = scene_target(p, scene, out, &sd);
  if (s) return s;
  if (!out->status || !out->rgba16f)
    return fail(GL_ERR_INVALID_ARGUMENT, "shadow refinement needs the status and RGBA16F planes");
  hipStream_t st = (hipStream_t)stream;
  s = join_side(p, st);  // the f64 passes of an overlapped launch may still write the planes
  if (s) return s;
  const size_t rays = size_t(sd->width) * sd->height;
  if (rays + 1 > p->edge_capacity) {
    if (p->d_edges) (void)hipFree(p->d_edges);
    p->d_edges = nullptr;
    p->edge_capacity = 0;
    GL_HIP(hipMalloc(&p->d_edges, (rays + 1) * sizeof(uint32_t)));
    p->edge_capacity = rays + 1;
  }
  s = upload_frame(p, {make_frame(scene, 0, sd->width * sd->height, *out)}, st);
  if (s) return s;
  hipLaunchKernelGGL(k_accel_reset, dim3(1), dim3(64), 0, st, (unsigned int*)nullptr, p->d_edges,
                     (QueueDev*)nullptr);
  plan_mark(p, 0, st);
  hipLaunchKernelGGL(k_shadow_classify, dim3(uint32_t((rays + 255) / 256)), dim3(256), 0, st,
                     (const uint32_t*)out->status, sd->width, sd->height, p->d_edges,
                     uint32_t(rays));
  GL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_shadow_refine, dim3(uint32_t(std::max(1, p->cu_count * 8))), dim3(64), 0, st,
                     p->d_scenes, p->d_frame, (const uint32_t*)p->d_edges, uint32_t(rays),
                     map_view(p));
  GL_HIP(hipGetLastError());
  plan_mark(p, 1, st);
  return GL_OK;
}

gl_status gl_trace_surface_footprint(gl_trace_plan* p, uint32_t scene, const uint32_t* pixels,
                                     size_t n, float* jacobian, uint32_t* meta, void* stream) {
  if (!p) return fail(GL_ERR_INVALID_ARGUMENT, "plan is required");
  if (scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
  if (!jacobian || !meta) return fail(GL_ERR_INVALID_ARGUMENT, "jacobian and meta are required");
  const SceneDev& sd = p->scenes[scene];
  const size_t rays = size_t(sd.width) * sd.height;
  if (!pixels && n != rays)
    return fail(GL_ERR_INVALID_ARGUMENT, "without a pixel list n must be the scene's pixel count");
  if (n > 0xffffffffull) return fail(GL_ERR_CAPACITY, "more than 2^32 footprints");
  if (sd.disk_in <= 0.0f) return fail(GL_ERR_INVALID_ARGUMENT, "the footprint needs a surface");
  if (n == 0) return GL_OK;
  hipStream_t st = (hipStream_t)stream;
  gl_status s = join_side(p, st);
  if (s) return s;
  gl_ray_outputs none{};
  none.count = rays;
  s = upload_frame(p, {make_frame(scene, 0, uint32_t(rays), none)}, st);
  if (s) return s;
  const uint32_t blocks = uint32_t(std::min<size_t>((n + kFootprintGroups - 1) / kFootprintGroups,
                                                    size_t(std::max(1, p->cu_count)) * 16));
  plan_mark(p, 0, st);
  hipLaunchKernelGGL(k_surface_footprint, dim3(blocks), dim3(64), 0, st, p->d_scenes, p->d_frame,
                     pixels, uint32_t(n), jacobian, meta);
  GL_HIP(hipGetLastError());
  plan_mark(p, 1, st);
  return GL_OK;
}

gl_status gl_trace_plan_shadow_edges(gl_trace_plan* p, uint64_t* count, void* stream) {
  if (!p || !count) return fail(GL_ERR_INVALID_ARGUMENT, "plan and count are required");
  *count = 0;
  if (!p->d_edges) return GL_OK;
  uint32_t c = 0;
  GL_HIP(hipMemcpyAsync(&c, p->d_edges, sizeof(uint32_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
  GL_HIP(hipStreamSynchronize((hipStream_t)stream));
  *count = c;
  return GL_OK;
}

}  // extern "C"
