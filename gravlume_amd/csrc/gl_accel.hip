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
// MI355X layout: a 64-lane wavefront owns an 8x8 block of nodes (the node
// pass) or one 8x8 tile (the tile pass), so the stencil is one LDS vector
// per wave and the accept decision is wave-uniform; node pixels keep their
// exact full-trace outputs (written by the node pass), so no quantized
// direction is ever reused. Refinement packs the four sub-samples of a
// pixel into a lane quad and reduces them with cross-lane shuffles.
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

__global__ void k_accel_reset(unsigned int* defer_count, uint32_t* edge_count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (defer_count) {
      defer_count[0] = 0;
      defer_count[1] = 0;
    }
    if (edge_count) edge_count[0] = 0;
  }
}

// Node pass: lane -> node (nx, ny) of an 8x8 node block, node -> pixel
// (4 nx, 4 ny). The node pixel's outputs are final; the node record keeps
// its class and unit escape direction for the tile pass.
__global__ void __launch_bounds__(64) k_sky_nodes(const SceneDev* __restrict__ scenes,
                                                  const FrameScene* __restrict__ frame,
                                                  uint32_t node_w, uint32_t node_h,
                                                  uint32_t blocks_x, float4* __restrict__ nodes,
                                                  MapView map, DeferBuf df) {
  const FrameScene& fsc = frame[0];
  const SceneDev& s = scenes[fsc.scene];
  SceneT<float> sc = scene_f32(s);
  hold_in_vgprs(sc);
  const uint32_t nx = (blockIdx.x % blocks_x) * 8u + (threadIdx.x & 7u);
  const uint32_t ny = (blockIdx.x / blocks_x) * 8u + (threadIdx.x >> 3);
  if (nx >= node_w || ny >= node_h) return;
  const uint32_t node = ny * node_w + nx;
  const uint32_t px = nx * 4u, py = ny * 4u;
  float4 rec = make_float4(0.0f, 0.0f, 0.0f, kNodeOther);
  if (px >= s.width || py >= s.height) {
    nodes[node] = rec;
    return;
  }
  const uint32_t idx = py * s.width + px;
  RayT<float> R;
  init_image_ray(s, sc, px, py, R);
  for (uint32_t it = 1; R.term == kTermActive; ++it) {
    if (wants_defer(sc, R) && try_defer(df, R, 0u, idx)) {
      nodes[node] = rec;  // finished in f64 later: not a stencil node
      return;
    }
    step<float>(sc, R, s.omega, reproject_now(sc, it), drift_now(it));
  }
  write_ray(R, fsc.out, idx, s, map);
  if (R.term == kTermHorizon) {
    rec.w = kNodeHorizon;
  } else if (R.term == kTermEscape && !(R.flags & kFlagTied)) {
    const unsigned long long n = fsc.out.count;
    const float* ob = fsc.out.observable;
    const float dx = ob[idx], dy = ob[n + idx], dz = ob[2 * n + idx];
    if (dx * dx + dy * dy + dz * dz > 0.5f) rec = make_float4(dx, dy, dz, kNodeEscape);
  }
  nodes[node] = rec;
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
// crossings, winding) are zero.
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

// Tile pass: one wavefront per 8x8 tile.
__global__ void __launch_bounds__(64) k_sky_tiles(const SceneDev* __restrict__ scenes,
                                                  const FrameScene* __restrict__ frame,
                                                  uint32_t tiles_x, uint32_t node_w,
                                                  const float4* __restrict__ nodes, float max_err2,
                                                  MapView map, DeferBuf df) {
  __shared__ float4 st[9];
  const FrameScene& fsc = frame[0];
  const SceneDev& s = scenes[fsc.scene];
  const uint32_t tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
  const uint32_t lane = threadIdx.x;
  if (lane < 9u) st[lane] = nodes[(2u * ty + lane / 3u) * node_w + 2u * tx + lane % 3u];
  __syncthreads();
  bool accept = true;
#pragma unroll
  for (int k = 0; k < 9; ++k) accept = accept && st[k].w == kNodeEscape;
  accept = accept && node_matches(st, 1, 0.5f, 0.0f, max_err2) &&
           node_matches(st, 3, 0.0f, 0.5f, max_err2) && node_matches(st, 4, 0.5f, 0.5f, max_err2) &&
           node_matches(st, 5, 1.0f, 0.5f, max_err2) && node_matches(st, 7, 0.5f, 1.0f, max_err2);
  const uint32_t lx = lane & 7u, ly = lane >> 3;
  const uint32_t px = tx * 8u + lx, py = ty * 8u + ly;
  if (px >= s.width || py >= s.height) return;
  if ((lx & 3u) == 0u && (ly & 3u) == 0u) return;  // a node pixel: already final
  const uint32_t idx = py * s.width + px;
  if (accept) {
    const uint32_t cx = min(lx >> 2, 1u), cy = min(ly >> 2, 1u);
    const float fu = float(lx - 4u * cx) * 0.25f, fv = float(ly - 4u * cy) * 0.25f;
    const uint32_t first = cy * 3u + cx;
    const float3 up = blend(st[first], st[first + 1], fu);
    const float3 lo = blend(st[first + 3], st[first + 4], fu);
    const float3 d = unit(make_float3(up.x + (lo.x - up.x) * fv, up.y + (lo.y - up.y) * fv,
                                      up.z + (lo.z - up.z) * fv));
    write_reconstructed(fsc.out, idx, s, map, d);
    return;
  }
  SceneT<float> sc = scene_f32(s);
  hold_in_vgprs(sc);
  RayT<float> R;
  init_image_ray(s, sc, px, py, R);
  for (uint32_t it = 1; R.term == kTermActive; ++it) {
    if (wants_defer(sc, R) && try_defer(df, R, 0u, idx)) return;
    step<float>(sc, R, s.omega, reproject_now(sc, it), drift_now(it));
  }
  write_ray(R, fsc.out, idx, s, map);
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
  const uint32_t tiles_x = (sd->width + 7u) / 8u, tiles_y = (sd->height + 7u) / 8u;
  const uint32_t node_w = 2u * tiles_x + 1u, node_h = 2u * tiles_y + 1u;
  const size_t nodes = size_t(node_w) * node_h;
  if (nodes > p->node_capacity) {
    if (p->d_nodes) (void)hipFree(p->d_nodes);
    p->d_nodes = nullptr;
    p->node_capacity = 0;
    GL_HIP(hipMalloc(&p->d_nodes, nodes * sizeof(float4)));
    p->node_capacity = nodes;
  }
  const bool f64 = (sd->options & kOptF64Horizon) != 0;
  if (f64) {
    s = ensure_defer(p, size_t(sd->width) * sd->height);
    if (s) return s;
  }
  s = upload_frame(p, {make_frame(scene, 0, sd->width * sd->height, *out)}, st);
  if (s) return s;
  hipLaunchKernelGGL(k_accel_reset, dim3(1), dim3(64), 0, st, p->d_defer_count, (uint32_t*)nullptr);
  DeferBuf df = defer_buf(p);
  if (!f64) df.rec = nullptr;
  const double half_pixel = double(p->host[scene].tan_half_fov) / double(sd->height);
  const float max_err = float(std::min(half_pixel, double(kStencilMaxError)));
  const uint32_t blocks_x = (node_w + 7u) / 8u, blocks_y = (node_h + 7u) / 8u;
  plan_mark(p, 0, st);
  hipLaunchKernelGGL(k_sky_nodes, dim3(blocks_x * blocks_y), dim3(64), 0, st, p->d_scenes,
                     p->d_frame, node_w, node_h, blocks_x, p->d_nodes, map_view(p), df);
  GL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sky_tiles, dim3(tiles_x * tiles_y), dim3(64), 0, st, p->d_scenes,
                     p->d_frame, tiles_x, node_w, (const float4*)p->d_nodes, max_err * max_err,
                     map_view(p), df);
  GL_HIP(hipGetLastError());
  plan_mark(p, 1, st);
  if (f64) return launch_finisher(p, st);
  return GL_OK;
}

gl_status gl_trace_refine_shadow(gl_trace_plan* p, uint32_t scene, const gl_ray_outputs* out,
                                 void* stream) {
  const SceneDev* sd = nullptr;
  gl_status s = scene_target(p, scene, out, &sd);
  if (s) return s;
  if (!out->status || !out->rgba16f)
    return fail(GL_ERR_INVALID_ARGUMENT, "shadow refinement needs the status and RGBA16F planes");
  hipStream_t st = (hipStream_t)stream;
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
  hipLaunchKernelGGL(k_accel_reset, dim3(1), dim3(64), 0, st, (unsigned int*)nullptr, p->d_edges);
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
