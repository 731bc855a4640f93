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
#include "gl_device.h"
#include "gl_plan.h"
#include "gl_scene.h"
#include "gl_types.h"
#include "trace_device.h"

namespace gl {

// ---------------------------------------------------------------------------
// Kernels

__global__ void __launch_bounds__(64) k_plain_image(const SceneDev* __restrict__ scenes,
                                                    const FrameScene* __restrict__ frame,
                                                    uint32_t tile_x0, uint32_t tile_y0,
                                                    uint32_t tiles_x, MapView map, DeferBuf df) {
  const FrameScene& fsc = frame[0];
  const SceneDev& s = scenes[fsc.scene];
  SceneT<float> sc = scene_f32(s);
  if (fsc.out.turning_radius) sc.options |= kOptTurningRadius;
  hold_in_vgprs(sc);
  uint32_t tx = tile_x0 + blockIdx.x % tiles_x;
  uint32_t ty = tile_y0 + blockIdx.x / tiles_x;
  uint32_t px = tx * 8u + (threadIdx.x & 7u);
  uint32_t py = ty * 8u + (threadIdx.x >> 3);
  if (px >= s.width || py >= s.height) return;
  uint32_t idx = py * s.width + px;
  RayT<float> R;
  init_image_ray(s, sc, px, py, R);
  uint32_t it = 1;
  bool deferred = false;
  trace_f32(sc, R, s.omega, df, 0u, idx, it, deferred);
  if (!deferred) write_ray(R, fsc.out, idx, s, map, pixel_source(px, py), refine_sink(df), 0u);
}

__global__ void __launch_bounds__(64) k_geodesic_batch(const SceneDev* __restrict__ scenes,
                                                       const FrameScene* __restrict__ frame,
                                                       const double* __restrict__ states,
                                                       const float* __restrict__ dirs,
                                                       uint32_t n, MapView map, DeferBuf df) {
  const FrameScene& fsc = frame[0];
  const SceneDev& s = scenes[fsc.scene];
  SceneT<float> sc = scene_f32(s);
  if (fsc.out.turning_radius) sc.options |= kOptTurningRadius;
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  RayT<float> R;
  init_state_ray(s, sc, states + 8 * size_t(i), dirs[i] >= 0.0f ? 1.0f : -1.0f, R);
  uint32_t it = 1;
  bool deferred = false;
  trace_f32(sc, R, s.omega, df, 0u, i, it, deferred);
  if (!deferred)
    write_ray(R, fsc.out, i, s, map, batch_source(states + 8 * size_t(i)), refine_sink(df), 0u);
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

// Lane refill: a wavefront steps its rays in rounds of `interval` attempts
// (rounded up to a multiple of 16); between rounds, lanes whose ray ended
// write it, and when at least `threshold` lanes are free they take the next
// items (global item indices from the cursor up to `total`) with one
// atomicAdd per wavefront (ballot / popc / shfl). The stepping round holds
// only the DP5 attempt and the f64-island test: a finished ray is written at
// the start of the next round, so the output path's registers and branches
// stay out of the stepping loop, whose exits are wavefront-uniform
// (ballots). For kSingle, sc and omega hold the one scene's constants;
// kBatchOnly: every item is a geodesic-batch request (no image path).
template <bool kSingle, bool kBatchOnly = false>
__device__ __forceinline__ void refill_rounds(const SceneDev* __restrict__ scenes,
                                              const FrameScene* __restrict__ frame,
                                              uint32_t n_frame, unsigned long long* cursor,
                                              unsigned long long total, MapView map, DeferBuf df,
                                              uint32_t interval, uint32_t threshold,
                                              SceneT<float>& sc, float& omega,
                                              uint32_t& lane_attempts, uint32_t& refills,
                                              uint32_t& wave_steps) {
  const uint32_t lane = __lane_id();
  RayT<float> R;
  R.term = kTermActive;
  bool active = false;     // a ray in flight
  bool done = false;       // a ray that ended and is not yet written
  bool may_defer = false;  // the ray may still go to the f64 finisher
  uint32_t fs = 0, idx = 0;
  // the ray's attempt clock (attempt number, and the same mod the
  // re-projection interval), as in the plain and whole-tile kernels: a ray's
  // arithmetic does not depend on the wavefront it was handed to
  uint32_t it = 1, phase = 1;
  bool drained = false;  // wave-uniform: the cursor has no items left
  for (;;) {
    if (done) {
      const FrameScene& f = frame[fs];
      const SceneDev& s = scenes[f.scene];
      write_ray(R, f.out, idx, s, map, frame_source(f, s, idx), refine_sink(df), fs);
      done = false;
    }
    if (!drained) {
      unsigned long long free_mask = __ballot(!active);
      uint32_t nfree = __popcll(free_mask);
      if (nfree >= threshold || nfree == 64u) {
        uint32_t leader = __ffsll((long long)free_mask) - 1;
        unsigned long long base = 0;
        if (lane == leader) base = atomicAdd(cursor, (unsigned long long)nfree);
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
              if (f.out.turning_radius) sc.options |= kOptTurningRadius;
              omega = s.omega;
            }
            uint32_t local = uint32_t(item - f.first_item);
            bool have = false;
            if (kBatchOnly || f.states) {  // geodesic batch request (items beyond rays pad to 64)
              if (local < f.rays) {
                idx = local;
                init_state_ray(s, sc, f.states + 8 * size_t(local),
                               f.dirs[local] >= 0.0f ? 1.0f : -1.0f, R);
                have = true;
              }
            } else if (!kBatchOnly) {
              uint32_t px, py;
              image_pixel(local, s.tiles_x, px, py);
              if (px < s.width && py < s.height) {
                idx = py * s.width + px;
                init_image_ray(s, sc, px, py, R);
                have = true;
              }
            }
            if (have) {
              active = true;
              may_defer = (sc.options & kOptF64Horizon) != 0u;
              it = 1;
              phase = sc.reproject_interval > 1u ? 1u : 0u;
              if (R.term != kTermActive) {
                write_ray(R, f.out, idx, s, map, frame_source(f, s, idx), refine_sink(df), fs);
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
    // Rounds of a multiple of 16 attempts between refills: every ray starts
    // at a round boundary, so the attempt clocks of a wavefront's rays stay
    // in phase modulo 16, and the drift sample (every 4th attempt) and the
    // re-projection (every 16th, the default) run for all lanes on the same
    // trips instead of some lane's on every trip.
    const uint32_t round = (max(interval, 1u) + 15u) & ~15u;
    for (uint32_t k = 0; k < round; ++k) {
      ++wave_steps;
      if (active) {
        ++lane_attempts;
        bool deferred = false;
        if (may_defer && R.r < sc.r_defer) {
          deferred = try_defer(df, R, fs, idx);
          may_defer = deferred;  // a full deferral buffer stays full: finish in f32
        }
        if (deferred) {
          active = false;
        } else {
          const uint32_t period = sc.reproject_interval;
          step<float>(sc, R, omega, period != 0u && phase == 0u, drift_now(it));
          ++it;
          phase = phase + 1u == period ? 0u : phase + 1u;
          if (R.term != kTermActive) {
            active = false;
            done = true;
          }
        }
      }
      if ((k & 15u) == 15u) {
        if (!drained && __popcll(__ballot(!active)) >= threshold) break;
      }
      if (drained && __ballot(active) == 0ull) break;
    }
  }
}

// kWaves: minimum waves per SIMD the register allocation must allow
// (the second __launch_bounds__ argument); the autotuner picks it.
template <bool kSingle, int kWaves>
__global__ void __launch_bounds__(256, kWaves) k_persistent(const SceneDev* __restrict__ scenes,
                                                    const FrameScene* __restrict__ frame,
                                                    uint32_t n_frame, QueueDev* queue, MapView map,
                                                    DeferBuf df, uint32_t interval,
                                                    uint32_t threshold) {
  const uint32_t lane = __lane_id();
  SceneT<float> sc;
  float omega = 0.0f;
  if (kSingle) {
    sc = scene_f32(scenes[frame[0].scene]);
    if (frame[0].out.turning_radius) sc.options |= kOptTurningRadius;
    omega = scenes[frame[0].scene].omega;
    hold_in_vgprs(sc);
  }
  uint32_t lane_attempts = 0;  // per lane: DP5 attempts made
  uint32_t refills = 0;
  uint32_t wave_steps = 0;  // wave attempt slots (lane slots / 64)
  refill_rounds<kSingle>(scenes, frame, n_frame, &queue->next, queue->total, map, df, interval,
                         threshold, sc, omega, lane_attempts, refills, wave_steps);
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
  // image tiles lie below batch_first; the multi-entry kernel refills lanes
  // from the geodesic-batch requests beyond it once the tiles are out
  const unsigned long long tile_end = kSingle ? total : queue->batch_first;
  const uint32_t lane = __lane_id();
  uint32_t lane_attempts = 0, tiles = 0, wave_steps = 0;
  bool cost_changed = false;
  for (;;) {
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&queue->next, 64ull);
    base = __shfl(base, 0);
    if (base >= tile_end) break;
    ++tiles;
    const unsigned long long item = base + lane;
    // the single-scene variant looks its frame slice up per tile as well:
    // with a constant slice the compiler keeps the scene's packed constants
    // live across the tile loop, registers the trace loop needs (it spilled)
    const uint32_t fs = find_frame_scene(frame, n_frame, item);
    const FrameScene& f = frame[fs];
    const SceneDev& s = scenes[f.scene];
    SceneT<float> sc = scene_f32(s);
    if (f.out.turning_radius) sc.options |= kOptTurningRadius;
    hold_in_vgprs(sc);
    const uint32_t local = uint32_t(item - f.first_item);
    uint32_t it = 0, idx = 0, tile = 0xffffffffu;
    bool have = false;
    RayT<float> R;
    if (f.states) {  // geodesic batch: 64 consecutive requests (padded to 64)
      if (local < f.rays) {
        idx = local;
        init_state_ray(s, sc, f.states + 8 * size_t(local), f.dirs[local] >= 0.0f ? 1.0f : -1.0f,
                       R);
        have = true;
      }
    } else {
      // the queue's tile slot -> the tile it hands out this launch
      tile = s.tile_order[local >> 6];
      uint32_t px, py;
      image_pixel((tile << 6) | (local & 63u), s.tiles_x, px, py);
      if (px < s.width && py < s.height) {
        idx = py * s.width + px;
        init_image_ray(s, sc, px, py, R);
        have = true;
      }
    }
    if (have) {
      bool deferred = false;
      it = 1;
      trace_f32(sc, R, s.omega, df, fs, idx, it, deferred);
      lane_attempts += it - 1;
      if (!deferred) write_ray(R, f.out, idx, s, map, frame_source(f, s, idx), refine_sink(df), fs);
    }
    uint32_t longest = it;  // wave iterations of this tile: its longest lane
    for (int off = 32; off > 0; off >>= 1) longest = max(longest, (uint32_t)__shfl_xor(longest, off));
    wave_steps += longest > 0u ? longest - 1u : 0u;
    if (lane == 0 && tile != 0xffffffffu && s.tile_cost[tile] != longest) {
      s.tile_cost[tile] = longest;
      cost_changed = true;
    }
  }
  if (!kSingle && tile_end < total) {
    // Geodesic-batch requests (step counts from tens to ~10^4 within one
    // batch, so whole 64-request tiles would leave most lanes idle): lane
    // refill in rounds of 16 attempts, refilling at 16 free lanes.
    SceneT<float> bsc;
    float bomega = 0.0f;
    refill_rounds<false, true>(scenes, frame, n_frame, &queue->batch_next, total, map, df, 16u, 16u,
                         bsc, bomega, lane_attempts, tiles, wave_steps);
  }
  for (int off = 32; off > 0; off >>= 1) lane_attempts += __shfl_xor(lane_attempts, off);
  if (lane == 0) {
    atomicAdd(&queue->attempts, (unsigned long long)lane_attempts);
    atomicAdd(&queue->slots, 64ull * wave_steps);
    atomicAdd(&queue->refills, (unsigned long long)tiles);
    if (cost_changed) atomicOr(&queue->order_dirty, 1ull);
  }
}

// The next whole-tile launch's tile order for each image scene of a frame:
// the tiles by decreasing cost of this launch (wave iterations of the tile's
// longest lane), so the queue hands out the longest tiles first and the
// launch ends on short ones (longest-processing-time-first). A counting sort
// over 16 octave buckets of the cost, one workgroup per scene: thread i takes
// tiles i, i + 1024, ... (coalesced loads, 32 in flight) and keeps its own
// counter per bucket in LDS (no two threads touch one counter), one
// block-wide scan of the counters gives every thread its output offsets, and
// a second pass writes the tiles in place; the order within a bucket is
// fixed by (tile mod 1024, tile / 1024). Which wavefront traces a tile, and
// when, leaves every output unchanged.
constexpr uint32_t kOrderThreads = 1024u, kOrderBuckets = 16u;

__device__ __forceinline__ uint32_t cost_bucket(uint32_t cost) {
  if (cost == 0u) return kOrderBuckets - 1u;
  const uint32_t octave = 31u - uint32_t(__clz(cost));  // floor(log2 cost)
  return kOrderBuckets - 1u - min(kOrderBuckets - 1u, octave);
}

__global__ void __launch_bounds__(kOrderThreads) k_tile_order(const SceneDev* __restrict__ scenes,
                                                              const FrameScene* __restrict__ frame,
                                                              const QueueDev* __restrict__ queue) {
  __shared__ uint32_t cnt[kOrderBuckets * kOrderThreads];  // bucket-major, 64 KiB
  __shared__ uint32_t part[kOrderThreads];
  // no tile of the launch changed its cost: the order stands (a static view
  // is sorted once, after the launch that first measured it)
  if (queue->order_dirty == 0ull) return;
  const FrameScene& f = frame[blockIdx.x];
  if (f.states) return;
  for (uint32_t b = 0; b < blockIdx.x; ++b)  // a scene listed twice is sorted once
    if (frame[b].scene == f.scene && !frame[b].states) return;
  const SceneDev& s = scenes[f.scene];
  const uint32_t n = s.tiles_x * ((s.height + 7u) / 8u);
  const uint32_t* cost = s.tile_cost;
  uint32_t* order = const_cast<uint32_t*>(s.tile_order);
  const uint32_t tid = threadIdx.x;
  constexpr uint32_t kBatch = 32u;
  const uint32_t stride = kOrderThreads * kBatch;
#pragma unroll
  for (uint32_t k = 0; k < kOrderBuckets; ++k) cnt[k * kOrderThreads + tid] = 0u;
  for (uint32_t t0 = tid; t0 < n; t0 += stride) {
    uint32_t c[kBatch];
#pragma unroll
    for (uint32_t j = 0; j < kBatch; ++j) {
      const uint32_t t = t0 + j * kOrderThreads;
      c[j] = t < n ? cost[t] : 0xffffffffu;
    }
#pragma unroll
    for (uint32_t j = 0; j < kBatch; ++j)
      if (t0 + j * kOrderThreads < n) cnt[cost_bucket(c[j]) * kOrderThreads + tid] += 1u;
  }
  __syncthreads();
  // exclusive scan of cnt in bucket-major order: thread i takes entries
  // [16 i, 16 i + 16), then the 1024 partial sums are scanned
  uint32_t local[kOrderBuckets], sum = 0u;
#pragma unroll
  for (uint32_t j = 0; j < kOrderBuckets; ++j) {
    const uint32_t v = cnt[kOrderBuckets * tid + j];
    local[j] = sum;
    sum += v;
  }
  part[tid] = sum;
  __syncthreads();
  for (uint32_t d = 1; d < kOrderThreads; d <<= 1) {  // inclusive scan of part
    const uint32_t add = tid >= d ? part[tid - d] : 0u;
    __syncthreads();
    part[tid] += add;
    __syncthreads();
  }
  const uint32_t base = part[tid] - sum;  // exclusive
#pragma unroll
  for (uint32_t j = 0; j < kOrderBuckets; ++j) cnt[kOrderBuckets * tid + j] = base + local[j];
  __syncthreads();
  for (uint32_t t0 = tid; t0 < n; t0 += stride) {
    uint32_t c[kBatch];
#pragma unroll
    for (uint32_t j = 0; j < kBatch; ++j) {
      const uint32_t t = t0 + j * kOrderThreads;
      c[j] = t < n ? cost[t] : 0xffffffffu;
    }
#pragma unroll
    for (uint32_t j = 0; j < kBatch; ++j) {
      const uint32_t t = t0 + j * kOrderThreads;
      if (t < n) {
        const uint32_t at = cost_bucket(c[j]) * kOrderThreads + tid;
        const uint32_t pos = cnt[at];
        cnt[at] = pos + 1u;
        order[pos] = t;
      }
    }
  }
}

// The f32 view of an f64 ray for the output writer: the position rounded
// with its f64 remainder as the compensation term, so the terminal
// observables are still formed from the f64 position.
__device__ __forceinline__ RayT<float> narrow_ray(const RayT<double>& R) {
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
  W.armed = R.armed;
  W.candidates = R.candidates;
  W.flags = R.flags;
  W.accepted = R.accepted;
  W.rejected = R.rejected;
  W.t = float(R.t - R.ct);
  W.ct = 0.0f;
  W.lam = float(R.lam - R.clam);
  W.clam = 0.0f;
  W.dir = R.dir;
  W.d_null = float(R.d_null);
  W.d_lz = float(R.d_lz);
  W.d_q = float(R.d_q);
  W.turnings = R.turnings;
  W.crossings = R.crossings;
  W.winding = R.winding;
  W.polar_side = R.polar_side;
  W.r_turn = float(R.r_turn);
  W.d_crit = float(R.d_crit);
  return W;
}

// The f64 continuation of a deferred f32 ray: the position and the
// coordinate time and affine parameter with their compensation folded in.
__device__ __forceinline__ void widen_deferred(const RayT<float>& Rf, RayT<double>& R) {
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
  R.r_turn = Rf.r_turn;
  R.d_crit = Rf.d_crit;
}

// The claim order of an f64 pass: its records by claim bucket, bucket 0
// first (defer_bucket, refine_bucket: the longest chains first), from the
// bucket counts taken when the records were written. Within a bucket the
// slots follow the atomics; the order changes which lane traces a ray and
// when, never its outputs. When every lane of the pass claims at most one
// record (as on a 1080p frame's ~30,000 deferred rays) all chains start at
// once whatever the order, and the write order, which keeps the rays of one
// tile together in a wavefront, is kept. Records are `stride` bytes with
// their bucket at `bucket_offset`; 256 records per round: a bucket
// histogram in LDS, one global atomic per nonempty bucket to reserve its
// slots, then the scatter.
__global__ void __launch_bounds__(256) k_claim_order(const unsigned char* __restrict__ rec,
                                                     uint32_t stride, uint32_t bucket_offset,
                                                     const unsigned int* __restrict__ count_n,
                                                     uint32_t capacity,
                                                     const unsigned int* __restrict__ hist,
                                                     unsigned int* fill, uint32_t* order,
                                                     uint32_t lanes) {
  __shared__ uint32_t base[kClaimBuckets], local[kClaimBuckets], start[kClaimBuckets];
  unsigned int n = *count_n;
  if (n > capacity) n = capacity;
  if (n <= lanes) return;
  const uint32_t tid = threadIdx.x;
  if (tid < kClaimBuckets) base[tid] = hist[tid];
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (uint32_t k = 0; k < kClaimBuckets; ++k) {
      const uint32_t c = base[k];
      base[k] = acc;
      acc += c;
    }
  }
  for (uint32_t first = blockIdx.x * 256u; first < n; first += gridDim.x * 256u) {
    if (tid < kClaimBuckets) local[tid] = 0u;
    __syncthreads();
    const uint32_t i = first + tid;
    const bool valid = i < n;
    uint32_t b = 0u;
    if (valid) {
      b = *reinterpret_cast<const uint32_t*>(rec + size_t(i) * stride + bucket_offset);
      if (b >= kClaimBuckets) b = kClaimBuckets - 1u;
    }
    const uint32_t rank = valid ? atomicAdd(&local[b], 1u) : 0u;
    __syncthreads();
    if (tid < kClaimBuckets && local[tid] != 0u)
      start[tid] = base[tid] + atomicAdd(fill + tid, local[tid]);
    __syncthreads();
    if (valid) {
      const uint32_t slot = start[b] + rank;
      if (slot < n) order[slot] = i;
    }
    __syncthreads();
  }
}

// The record behind claim i (< n): the write order when the pass's lanes
// take one record each, else the ordering pass's permutation of [0, n) (an
// entry outside it falls back to record i).
__device__ __forceinline__ uint32_t claimed_record(const uint32_t* order, uint32_t i, uint32_t n) {
  if (n <= gridDim.x * blockDim.x) return i;
  const uint32_t r = order[i];
  return r < n ? r : i;
}

// Rays that entered the f64 island finish with f64 DP5 steps. The lanes
// refill one ray at a time from a claim cursor (df.count[1], through the
// claim order) inside one loop: a lane whose ray ends writes it and takes
// the next record at once, so a long f64 chain occupies its own lane only
// and the other lanes of its wavefront keep stepping (f64 chains range from
// one attempt to thousands).
__global__ void __launch_bounds__(64) k_finish_f64(const SceneDev* __restrict__ scenes,
                                                   const FrameScene* __restrict__ frame,
                                                   MapView map, DeferBuf df) {
  unsigned int n = *df.count;
  if (n > df.capacity) n = df.capacity;
  RayT<double> R;
  SceneT<double> sd;
  double omega = 0.0;
  uint32_t fs = 0, idx = 0, it = 1;
  unsigned int i = atomicAdd(df.count + 1, 1u);
  bool live = i < n;
  if (live) {
    const DeferRecord& rec = df.rec[claimed_record(df.order, i, n)];
    fs = rec.frame_scene;
    idx = rec.index;
    const SceneDev& s = scenes[frame[fs].scene];
    sd = scene_f64(s);
    if (frame[fs].out.turning_radius) sd.options |= kOptTurningRadius;
    omega = double(s.omega);
    widen_deferred(rec.ray, R);
    if (!refresh_derivative<double>(sd, R)) {
      R.term = kTermFailure;
      R.flags |= kFlagNonFinite;
    }
  }
  while (live) {
    if (R.term == kTermActive) {
      step<double>(sd, R, omega, false, drift_now(it));
      ++it;
    }
    if (R.term != kTermActive) {
      const FrameScene& f = frame[fs];
      const SceneDev& s = scenes[f.scene];
      write_ray(narrow_ray(R), f.out, idx, s, map, frame_source(f, s, idx), refine_sink(df), fs);
      i = atomicAdd(df.count + 1, 1u);
      live = i < n;
      if (live) {
        const DeferRecord& rec = df.rec[claimed_record(df.order, i, n)];
        fs = rec.frame_scene;
        idx = rec.index;
        const SceneDev& s2 = scenes[frame[fs].scene];
        sd = scene_f64(s2);
        if (frame[fs].out.turning_radius) sd.options |= kOptTurningRadius;
        omega = double(s2.omega);
        widen_deferred(rec.ray, R);
        it = 1;
        if (!refresh_derivative<double>(sd, R)) {
          R.term = kTermFailure;
          R.flags |= kFlagNonFinite;
        }
      }
    }
  }
}

// The exact f64 start of a refine record's ray, with the scene's refine
// tolerances.
__device__ __forceinline__ void start_refined(const SceneDev* __restrict__ scenes,
                                              const FrameScene* __restrict__ frame,
                                              const RefineRecord& rec, SceneT<double>& sd,
                                              RayT<double>& R, double& omega) {
  const FrameScene& f = frame[rec.frame_scene];
  const SceneDev& s = scenes[f.scene];
  sd = scene_f64(s);
  sd.rtol = double(s.refine_rtol);
  sd.atol = double(s.refine_atol);
  if (f.out.turning_radius) sd.options |= kOptTurningRadius;
  omega = double(s.omega);
  const RaySource src = frame_source(f, s, rec.index);
  double y[8];
  source_state(s, src.state8, src.px, src.py, y);
  const float dir = f.states ? (f.dirs[rec.index] >= 0.0f ? 1.0f : -1.0f) : -1.0f;
  const uint32_t side = f.states ? (y[3] > 0.0 ? 2u : (y[3] < 0.0 ? 0u : 1u)) : s.polar_side;
  init_ray<double>(sd, R, y[1], y[2], y[3], 0.0, 0.0, 0.0, y[5], y[6], y[7], -y[4], dir, side);
  R.h = double(dir) * double(s.h_init);
  apply_start(R, start_test(s, y, true));  // an exact start: no rounding bits
}

// The f64 refine pass (GL_OPT_NEAR_CRITICAL_REFINE): every near-critical
// ray of the launch is traced again from its exact f64 initial state, in
// f64 throughout at the scene's refine tolerances, and written with the
// near-critical test of that trace (an f64 start and f64 drifts). One ray
// per lane from a claim cursor (df.count[3]), refilled as in the finisher.
__global__ void __launch_bounds__(64) k_refine_f64(const SceneDev* __restrict__ scenes,
                                                   const FrameScene* __restrict__ frame,
                                                   MapView map, DeferBuf df) {
  unsigned int n = df.count[2];
  if (n > df.refine_capacity) n = df.refine_capacity;
  RayT<double> R;
  SceneT<double> sd;
  double omega = 0.0;
  RefineRecord rec{0u, 0u, 0u};
  uint32_t it = 1;
  unsigned int i = atomicAdd(df.count + 3, 1u);
  bool live = i < n;
  if (live) {
    rec = df.refine[claimed_record(df.refine_order, i, n)];
    start_refined(scenes, frame, rec, sd, R, omega);
  }
  while (live) {
    if (R.term == kTermActive) {
      step<double>(sd, R, omega, false, drift_now(it));
      ++it;
    }
    if (R.term != kTermActive) {
      const FrameScene& f = frame[rec.frame_scene];
      const SceneDev& s = scenes[f.scene];
      write_ray(narrow_ray(R), f.out, rec.index, s, map, frame_source(f, s, rec.index),
                no_refine(), rec.frame_scene, true);
      i = atomicAdd(df.count + 3, 1u);
      live = i < n;
      if (live) {
        rec = df.refine[claimed_record(df.refine_order, i, n)];
        start_refined(scenes, frame, rec, sd, R, omega);
        it = 1;
      }
    }
  }
}

// The exact f64 start of geodesic-batch request i.
__device__ __forceinline__ void start_batch_f64(const SceneDev& s, const SceneT<double>& sd,
                                                const FrameScene& f, uint32_t i,
                                                RayT<double>& R) {
  double y[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) y[k] = f.states[8 * size_t(i) + k];
  const float dir = f.dirs[i] >= 0.0f ? 1.0f : -1.0f;
  const uint32_t side = y[3] > 0.0 ? 2u : (y[3] < 0.0 ? 0u : 1u);
  init_ray<double>(sd, R, y[1], y[2], y[3], 0.0, 0.0, 0.0, y[5], y[6], y[7], -y[4], dir, side);
  R.h = double(dir) * double(s.h_init);
  apply_start(R, start_test(s, y, true));  // an exact start: no rounding bits
}

// The f64 geodesic batch (gl_geodesic_batch_trace_f64): every request
// traced in f64 throughout from its exact initial state at the caller's
// tolerances, the reference batch's arithmetic (batch.rs:31-43 over
// geodesic.rs:129-212). Lanes refill from a claim cursor as in the
// finisher. Written as refined rays (GL_STATUS_REFINED).
__global__ void __launch_bounds__(64) k_batch_f64(const SceneDev* __restrict__ scenes,
                                                  const FrameScene* __restrict__ frame,
                                                  MapView map, unsigned int* cursor, uint32_t n,
                                                  double rtol, double atol) {
  const FrameScene& f = frame[0];
  const SceneDev& s = scenes[f.scene];
  SceneT<double> sd = scene_f64(s);
  sd.rtol = rtol;
  sd.atol = atol;
  if (f.out.turning_radius) sd.options |= kOptTurningRadius;
  const double omega = double(s.omega);
  RayT<double> R;
  uint32_t it = 1;
  unsigned int i = atomicAdd(cursor, 1u);
  bool live = i < n;
  if (live) start_batch_f64(s, sd, f, i, R);
  while (live) {
    if (R.term == kTermActive) {
      step<double>(sd, R, omega, false, drift_now(it));
      ++it;
    }
    if (R.term != kTermActive) {
      write_ray(narrow_ray(R), f.out, i, s, map, batch_source(f.states + 8 * size_t(i)),
                no_refine(), 0u, true);
      i = atomicAdd(cursor, 1u);
      live = i < n;
      if (live) {
        start_batch_f64(s, sd, f, i, R);
        it = 1;
      }
    }
  }
}

__global__ void k_queue_reset(QueueDev* q, unsigned long long total,
                              unsigned long long batch_first, unsigned int* defer_count) {
  if (blockIdx.x != 0) return;
  if (defer_count) clear_counters(defer_count);  // deferral, refine and ordering counters
  if (threadIdx.x == 0) {
    q->next = 0;
    q->total = total;
    q->batch_first = batch_first;
    q->batch_next = batch_first;
    q->refills = 0;
    q->attempts = 0;
    q->slots = 0;
    q->order_dirty = 0;
  }
}

}  // namespace gl

// ===========================================================================
// Host side

using namespace gl;

namespace gl {
thread_local std::string g_last_error;
gl_status fail(gl_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}
}  // namespace gl

namespace gl {

gl_status hip_fail(hipError_t e, const char* what) {
  return fail(GL_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

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
  p.rgba16f = o.rgba16f;
  p.event_residual = o.event_residual;
  p.turning_radius = o.turning_radius;
  p.azimuth_advance = o.azimuth_advance;
  p.count = o.count;
  return p;
}

}  // namespace gl

struct gl_frame_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  FrameScene* d_table = nullptr;
  gl::Workspace ws;  // the graph's own deferral / refine buffers and counters
  gl_trace_plan* plan = nullptr;
};

namespace gl {

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

gl_status ensure_workspace(Workspace& w, size_t rays, bool f64, bool refine) {
  // At most one record per ray, capped at 4M rays (~0.8 GB of deferral
  // records); beyond the cap a ray keeps stepping in f32 (deferral) or is
  // written as near-critical and Uncertain (refine).
  const size_t want = std::min<size_t>(std::max<size_t>(rays, 1), size_t(1) << 22);
  if (!w.counters) {
    GL_HIP(hipMalloc(&w.counters, kCounterWords * sizeof(unsigned int)));
    GL_HIP(hipMemset(w.counters, 0, kCounterWords * sizeof(unsigned int)));
  }
  if (f64 && (want > w.defer_capacity || !w.defer || !w.order)) {
    if (w.defer) (void)hipFree(w.defer);
    if (w.order) (void)hipFree(w.order);
    w.defer = nullptr;
    w.order = nullptr;
    w.defer_capacity = 0;
    GL_HIP(hipMalloc(&w.defer, want * sizeof(DeferRecord)));
    GL_HIP(hipMalloc(&w.order, want * sizeof(uint32_t)));
    w.defer_capacity = uint32_t(want);
  }
  if (refine && (want > w.refine_capacity || !w.refine || !w.refine_order)) {
    if (w.refine) (void)hipFree(w.refine);
    if (w.refine_order) (void)hipFree(w.refine_order);
    w.refine = nullptr;
    w.refine_order = nullptr;
    w.refine_capacity = 0;
    GL_HIP(hipMalloc(&w.refine, want * sizeof(RefineRecord)));
    GL_HIP(hipMalloc(&w.refine_order, want * sizeof(uint32_t)));
    w.refine_capacity = uint32_t(want);
  }
  return GL_OK;
}

void free_workspace(Workspace& w) {
  if (w.defer) (void)hipFree(w.defer);
  if (w.order) (void)hipFree(w.order);
  if (w.refine) (void)hipFree(w.refine);
  if (w.refine_order) (void)hipFree(w.refine_order);
  if (w.counters) (void)hipFree(w.counters);
  w = Workspace();
}

DeferBuf workspace_view(const Workspace& w, bool f64, bool refine) {
  DeferBuf d;
  d.rec = f64 ? w.defer : nullptr;
  d.order = f64 ? w.order : nullptr;
  d.count = w.counters;
  d.capacity = f64 ? w.defer_capacity : 0u;
  d.refine = refine ? w.refine : nullptr;
  d.refine_order = refine ? w.refine_order : nullptr;
  d.refine_capacity = refine ? w.refine_capacity : 0u;
  return d;
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

}  // namespace gl

// Event 0 opens a new timed launch (main kernel start), 1 closes the main
// kernel, 2 and 3 bracket the f64 finisher of the same launch.
void gl::plan_mark(gl_trace_plan* p, int which, hipStream_t st) {
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

namespace gl {

void mark(gl_trace_plan* p, int which, hipStream_t st) { gl::plan_mark(p, which, st); }

MapView map_view(const gl_trace_plan* p) {
  MapView m;
  m.rgb = p->d_map;
  m.w = p->map_w;
  m.h = p->map_h;
  return m;
}

bool scene_f64_horizon(const SceneDev& s) { return (s.options & kOptF64Horizon) != 0; }
bool scene_refine(const SceneDev& s) {
  return (s.options & kOptRefine) != 0 && s.near_critical_factor > 0.0f && s.r_horizon > 0.0f;
}

bool any_f64_horizon(const gl_trace_plan* p) {
  for (const auto& s : p->scenes)
    if (scene_f64_horizon(s)) return true;
  return false;
}

bool any_refine(const gl_trace_plan* p) {
  for (const auto& s : p->scenes)
    if (scene_refine(s)) return true;
  return false;
}

gl_status check_outputs(const gl_ray_outputs* o, size_t rays) {
  if (!o) return fail(GL_ERR_INVALID_ARGUMENT, "outputs must not be NULL");
  if (o->count < rays) return fail(GL_ERR_INVALID_ARGUMENT, "output planes hold fewer rays than the scene");
  if (reinterpret_cast<uintptr_t>(o->rgba16f) % 8u != 0u)
    return fail(GL_ERR_INVALID_ARGUMENT, "the RGBA16F plane must be 8-byte aligned");
  return GL_OK;
}

// The process-wide band table (gl_spectral.cpp), built on first use.
const std::vector<float>* blackbody_table() {
  static const std::vector<float> table = [] {
    std::vector<float> t(size_t(GL_BLACKBODY_LUT_ENTRIES) * 4);
    if (gl_blackbody_lut(t.data(), t.size()) != GL_OK) t.clear();
    return t;
  }();
  return table.empty() ? nullptr : &table;
}

// The claim orders and the f64 passes of a launch on `st`: the finisher
// over however many rays were deferred, then the refine pass over however
// many were near-critical (all counts read on the device).
gl_status launch_f64_kernels(const gl_trace_plan* p, const FrameScene* table, const DeferBuf& df,
                             hipStream_t st) {
  const dim3 grid(std::max(1, p->cu_count * 4)), order_grid(std::max(1, p->cu_count));
  const uint32_t lanes = grid.x * 64u;
  if (df.rec) {
    hipLaunchKernelGGL(k_claim_order, order_grid, dim3(256), 0, st,
                       reinterpret_cast<const unsigned char*>(df.rec), uint32_t(sizeof(DeferRecord)),
                       uint32_t(offsetof(DeferRecord, bucket)), df.count, df.capacity,
                       df.count + kDeferHist, df.count + kDeferFill, df.order, lanes);
    GL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finish_f64, grid, dim3(64), 0, st, p->d_scenes, table, map_view(p), df);
    GL_HIP(hipGetLastError());
  }
  if (df.refine) {
    hipLaunchKernelGGL(k_claim_order, order_grid, dim3(256), 0, st,
                       reinterpret_cast<const unsigned char*>(df.refine),
                       uint32_t(sizeof(RefineRecord)), uint32_t(offsetof(RefineRecord, bucket)),
                       df.count + 2, df.refine_capacity, df.count + kRefineHist,
                       df.count + kRefineFill, df.refine_order, lanes);
    GL_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_refine_f64, grid, dim3(64), 0, st, p->d_scenes, table, map_view(p), df);
    GL_HIP(hipGetLastError());
  }
  return GL_OK;
}

// The finisher over however many rays were deferred, then the refine pass
// over however many were near-critical (both counts read on the device).
// Timing events 2 and 3 bracket both.
gl_status launch_f64_passes(gl_trace_plan* p, const FrameScene* table, const DeferBuf& df,
                            hipStream_t st) {
  if (!df.rec && !df.refine) return GL_OK;
  mark(p, 2, st);
  gl_status s = launch_f64_kernels(p, table, df, st);
  if (s) return s;
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

// Frame table of a mixed work list: image scenes (all pixels, 8x8-tile
// ordered) and geodesic batches (requests padded to a multiple of 64 so that
// every 64-item chunk lies inside one entry). The image entries come first,
// in the list's order, then the batches: in whole-tile mode the image tiles
// are handed out whole and the batch requests from *batch_first on lane by
// lane. Which entry a ray belongs to changes none of its outputs.
gl_status build_work_table(gl_trace_plan* p, const gl_work_item* items, size_t n,
                           std::vector<FrameScene>* table, unsigned long long* total_items,
                           size_t* total_rays, unsigned long long* batch_first) {
  if (!items || n == 0) return fail(GL_ERR_INVALID_ARGUMENT, "no work items");
  table->resize(n);
  unsigned long long total = 0;
  size_t rays_sum = 0;
  std::vector<size_t> order;
  order.reserve(n);
  for (size_t i = 0; i < n; ++i)
    if (items[i].kind != GL_WORK_GEODESIC_BATCH) order.push_back(i);
  for (size_t i = 0; i < n; ++i)
    if (items[i].kind == GL_WORK_GEODESIC_BATCH) order.push_back(i);
  *batch_first = ~0ull;
  for (size_t k = 0; k < n; ++k) {
    const size_t i = order[k];
    const gl_work_item& w = items[i];
    if (w.kind == GL_WORK_GEODESIC_BATCH && *batch_first == ~0ull) *batch_first = total;
    if (w.scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
    const SceneDev& sd = p->scenes[w.scene];
    size_t rays = 0;
    unsigned long long cnt = 0;
    if (w.kind == GL_WORK_IMAGE) {
      rays = size_t(sd.width) * sd.height;
      cnt = 64ull * sd.tiles_x * ((sd.height + 7) / 8);
    } else if (w.kind == GL_WORK_GEODESIC_BATCH) {
      if (w.count > 0 && (!w.states8 || !w.directions))
        return fail(GL_ERR_INVALID_ARGUMENT, "geodesic batch item needs states and directions");
      if (w.count > 0xffffffffull) return fail(GL_ERR_CAPACITY, "more than 2^32 requests");
      rays = w.count;
      cnt = (w.count + 63ull) / 64ull * 64ull;
    } else {
      return fail(GL_ERR_INVALID_ARGUMENT, "unknown work item kind");
    }
    gl_status s = check_outputs(&w.outputs, rays);
    if (s) return s;
    if (total + cnt > 0xffffffffull) return fail(GL_ERR_CAPACITY, "work list exceeds 2^32 items");
    FrameScene f = make_frame(w.scene, uint32_t(total), uint32_t(rays), w.outputs);
    if (w.kind == GL_WORK_GEODESIC_BATCH) {
      f.states = w.states8;
      f.dirs = w.directions;
    }
    (*table)[k] = f;
    total += cnt;
    rays_sum += rays;
  }
  if (*batch_first == ~0ull) *batch_first = total;
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
// steps or the near-critical refine) the f64 passes, all on `st`, reading
// the frame table d_table and using workspace `w` (the plan's, or a frame
// graph's own).
gl_status launch_persistent(gl_trace_plan* p, const FrameScene* d_table, size_t n_table,
                            unsigned long long total, size_t total_rays, hipStream_t st,
                            const gl_persistent_config& c, Workspace& w_in,
                            unsigned long long batch_first) {
  const bool f64 = any_f64_horizon(p), refine = any_refine(p);
  // f64 overlap: an eager launch (the plan's workspace, not being captured)
  // alternates between two workspaces and frame-table copies, and runs its
  // f64 passes on the plan's second stream
  const bool ov = p->overlap && p->side && !p->capturing && &w_in == &p->ws && (f64 || refine);
  const int k = p->ov_parity;
  if (ov && p->fin_recorded[k]) {
    // launch N-2's f64 passes read this workspace and frame copy: the host
    // waits for them before it may grow or rewrite either
    GL_HIP(hipEventSynchronize(p->ev_fin[k]));
    p->fin_recorded[k] = false;
  }
  if (!ov && &w_in == &p->ws) {
    // a serial launch on the plan's workspace: no earlier pass may still
    // read it on the second stream
    gl_status d = drain_side(p);
    if (d) return d;
  }
  Workspace& w = ov && k ? p->ws_alt : w_in;
  gl_status s = ensure_workspace(w, total_rays, f64, refine);
  if (s) return s;
  const FrameScene* f64_table = d_table;
  if (ov) {
    if (p->frame_ov_capacity[k] < n_table) {
      if (p->d_frame_ov[k]) GL_HIP(hipFree(p->d_frame_ov[k]));
      p->d_frame_ov[k] = nullptr;
      p->frame_ov_capacity[k] = 0;
      GL_HIP(hipMalloc(&p->d_frame_ov[k], n_table * sizeof(FrameScene)));
      p->frame_ov_capacity[k] = n_table;
    }
    GL_HIP(hipMemcpyAsync(p->d_frame_ov[k], d_table, n_table * sizeof(FrameScene),
                          hipMemcpyDeviceToDevice, st));
    f64_table = p->d_frame_ov[k];
  }
  // the lane-by-lane batch phase belongs to the multi-entry whole-tile kernel
  if (n_table == 1 || c.refill_threshold < 64u) batch_first = total;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, p->d_queue, total, batch_first,
                     w.counters);
  GL_HIP(hipGetLastError());
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
  const DeferBuf df = workspace_view(w, f64, refine);
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
  if (ov) {
    GL_HIP(hipEventRecord(p->ev_main[k], st));
    GL_HIP(hipStreamWaitEvent(p->side, p->ev_main[k], 0));
    s = launch_f64_passes(p, f64_table, df, p->side);
    if (s) {
      // nothing of a failed launch may stay in flight unrecorded
      (void)hipStreamSynchronize(p->side);
      return s;
    }
    GL_HIP(hipEventRecord(p->ev_fin[k], p->side));
    p->fin_recorded[k] = true;
    p->fin_unjoined[k] = true;
    // the previous launch's passes ran beside this trace: join them
    if (p->fin_unjoined[1 - k]) {
      GL_HIP(hipStreamWaitEvent(st, p->ev_fin[1 - k], 0));
      p->fin_unjoined[1 - k] = false;
    }
    p->ov_parity = 1 - k;
  } else {
    s = launch_f64_passes(p, d_table, df, st);
    if (s) return s;
  }
  // the next whole-tile launch hands out this launch's longest tiles first
  if (tiles) return launch_tile_order(p, d_table, n_table, st);
  return GL_OK;
}

// f64 overlap: `st` waits for the passes still on the plan's second stream.
gl_status join_side(gl_trace_plan* p, hipStream_t st) {
  for (int k = 0; k < 2; ++k)
    if (p->fin_unjoined[k]) {
      GL_HIP(hipStreamWaitEvent(st, p->ev_fin[k], 0));
      p->fin_unjoined[k] = false;
    }
  return GL_OK;
}

// f64 overlap: the host waits until the plan's second stream has finished
// every pass, so that an entry point that reallocates or rewrites the
// workspaces, or captures a graph, sees none of them in flight.
gl_status drain_side(gl_trace_plan* p) {
  for (int k = 0; k < 2; ++k) {
    if (p->fin_recorded[k]) GL_HIP(hipEventSynchronize(p->ev_fin[k]));
    p->fin_recorded[k] = false;
    p->fin_unjoined[k] = false;
  }
  return GL_OK;
}

gl_status launch_tile_order(gl_trace_plan* p, const FrameScene* d_table, size_t n_table,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_tile_order, dim3(uint32_t(n_table)), dim3(kOrderThreads), 0, st,
                     p->d_scenes, d_table, (const QueueDev*)p->d_queue);
  GL_HIP(hipGetLastError());
  return GL_OK;
}

gl_status check_config(const gl_persistent_config& c) {
  if (c.block_size < 64 || c.block_size > 256 || c.block_size % 64 != 0)
    return fail(GL_ERR_INVALID_ARGUMENT, "block_size must be 64, 128, 192 or 256");
  if (c.blocks_per_cu == 0 || c.blocks_per_cu > 32)
    return fail(GL_ERR_INVALID_ARGUMENT, "blocks_per_cu must lie in [1, 32]");
  if (c.refill_interval == 0 || c.refill_interval > 4096)
    return fail(GL_ERR_INVALID_ARGUMENT, "refill_interval must lie in [1, 4096]");
  if (c.refill_threshold == 0 || c.refill_threshold > 64)
    return fail(GL_ERR_INVALID_ARGUMENT, "refill_threshold must lie in [1, 64]");
  if (c.waves_per_simd < 2 || c.waves_per_simd > 4)
    return fail(GL_ERR_INVALID_ARGUMENT, "waves_per_simd must be 2, 3 or 4");
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
  return launch_persistent(p, p->d_frame, table.size(), total, rays, st, p->pcfg, p->ws, total);
}

}  // namespace gl

extern "C" {

const char* gl_last_error(void) { return gl::g_last_error.c_str(); }

const char* gl_version(void) { return "gravlume-mi355x 0.1 (gfx950, f32 DP5(4) trace)"; }

gl_status gl_scene_validate(const gl_scene_desc* scene) {
  if (!scene) return fail(GL_ERR_INVALID_ARGUMENT, "scene is required");
  SceneDev sd;
  SceneHost hs;
  std::string err;
  int st = prepare_scene(*scene, &sd, &hs, &err);
  if (st != GL_OK) return fail(gl_status(st), err);
  return GL_OK;
}

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
  bool blackbody = false;
  for (const auto& s : p->scenes) blackbody = blackbody || s.emission_model == GL_EMISSION_BLACKBODY;
  if (blackbody) {
    const std::vector<float>* lut = blackbody_table();
    if (!lut) {
      gl_trace_plan_destroy(p);
      return fail(GL_ERR_NOT_REPRESENTABLE, "blackbody band table could not be built");
    }
    const size_t bytes = lut->size() * sizeof(float);
    if ((e = hipMalloc(&p->d_lut, bytes)) != hipSuccess) return cleanup(e, "hipMalloc band table");
    if ((e = hipMemcpy(p->d_lut, lut->data(), bytes, hipMemcpyHostToDevice)) != hipSuccess)
      return cleanup(e, "upload band table");
    for (auto& s : p->scenes) s.lut = p->d_lut;
  }
  {
    // tile orders (identity to start) and tile costs of every scene
    size_t total_tiles = 0;
    for (const auto& s : p->scenes) total_tiles += size_t(s.tiles_x) * ((s.height + 7u) / 8u);
    std::vector<uint32_t> init(2 * total_tiles, 0u);
    size_t off = 0;
    for (const auto& s : p->scenes) {
      const size_t n = size_t(s.tiles_x) * ((s.height + 7u) / 8u);
      for (size_t t = 0; t < n; ++t) init[off + t] = uint32_t(t);
      off += n;
    }
    if ((e = hipMalloc(&p->d_tiles, std::max<size_t>(init.size(), 1) * sizeof(uint32_t))) != hipSuccess)
      return cleanup(e, "hipMalloc tile order");
    if (!init.empty() &&
        (e = hipMemcpy(p->d_tiles, init.data(), init.size() * sizeof(uint32_t), hipMemcpyHostToDevice)) != hipSuccess)
      return cleanup(e, "upload tile order");
    off = 0;
    for (auto& s : p->scenes) {
      const size_t n = size_t(s.tiles_x) * ((s.height + 7u) / 8u);
      s.tile_order = p->d_tiles + off;
      s.tile_cost = p->d_tiles + total_tiles + off;
      off += n;
    }
  }
  if ((e = hipMalloc(&p->d_scenes, n_scenes * sizeof(SceneDev))) != hipSuccess) return cleanup(e, "hipMalloc scenes");
  if ((e = hipMemcpy(p->d_scenes, p->scenes.data(), n_scenes * sizeof(SceneDev), hipMemcpyHostToDevice)) != hipSuccess)
    return cleanup(e, "upload scenes");
  if ((e = hipMalloc(&p->d_queue, sizeof(QueueDev))) != hipSuccess) return cleanup(e, "hipMalloc queue");
  if ((e = hipMemset(p->d_queue, 0, sizeof(QueueDev))) != hipSuccess) return cleanup(e, "clear queue");
  if ((e = hipMalloc(&p->ws.counters, kCounterWords * sizeof(unsigned int))) != hipSuccess)
    return cleanup(e, "hipMalloc counters");
  if ((e = hipMemset(p->ws.counters, 0, kCounterWords * sizeof(unsigned int))) != hipSuccess)
    return cleanup(e, "clear counters");
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
  if (p->side) {
    (void)hipStreamSynchronize(p->side);
    (void)hipStreamDestroy(p->side);
  }
  for (int k = 0; k < 2; ++k) {
    if (p->ev_main[k]) (void)hipEventDestroy(p->ev_main[k]);
    if (p->ev_fin[k]) (void)hipEventDestroy(p->ev_fin[k]);
    if (p->d_frame_ov[k]) (void)hipFree(p->d_frame_ov[k]);
  }
  free_workspace(p->ws_alt);
  if (p->d_scenes) (void)hipFree(p->d_scenes);
  if (p->d_map) (void)hipFree(p->d_map);
  if (p->d_lut) (void)hipFree(p->d_lut);
  if (p->d_nodes) (void)hipFree(p->d_nodes);
  if (p->d_ready) (void)hipFree(p->d_ready);
  if (p->d_edges) (void)hipFree(p->d_edges);
  if (p->d_queue) (void)hipFree(p->d_queue);
  if (p->d_frame) (void)hipFree(p->d_frame);
  if (p->h_frame) (void)hipHostFree(p->h_frame);
  free_workspace(p->ws);
  if (p->d_tiles) (void)hipFree(p->d_tiles);
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
  const bool f64 = scene_f64_horizon(sd), refine = scene_refine(sd);
  s = drain_side(p);
  if (s) return s;
  s = ensure_workspace(p->ws, size_t(region->tiles_x) * region->tiles_y * 64, f64, refine);
  if (s) return s;
  s = upload_frame(p, {make_frame(region->scene, 0, sd.width * sd.height, *out)}, st);
  if (s) return s;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, p->d_queue, 0ull, 0ull, p->ws.counters);
  const DeferBuf df = workspace_view(p->ws, f64, refine);
  size_t blocks = size_t(region->tiles_x) * region->tiles_y;
  mark(p, 0, st);
  hipLaunchKernelGGL(k_plain_image, dim3(uint32_t(blocks)), dim3(64), 0, st, p->d_scenes,
                     p->d_frame, region->tile_x, region->tile_y, region->tiles_x, map_view(p), df);
  GL_HIP(hipGetLastError());
  mark(p, 1, st);
  return launch_f64_passes(p, p->d_frame, df, st);
}

gl_status gl_trace_encode_persistent(gl_trace_plan* p, const uint32_t* scenes, size_t n_scenes,
                                     const gl_ray_outputs* outputs, void* stream) {
  if (!p) return fail(GL_ERR_INVALID_ARGUMENT, "plan is required");
  return persistent_frame(p, scenes, n_scenes, outputs, (hipStream_t)stream);
}

gl_status gl_trace_plan_set_f64_overlap(gl_trace_plan* p, int enable) {
  if (!p) return fail(GL_ERR_INVALID_ARGUMENT, "plan is required");
  if (!enable) {
    gl_status s = drain_side(p);
    p->overlap = false;
    return s;
  }
  for (int k = 0; k < 2; ++k) {
    if (!p->ev_main[k]) GL_HIP(hipEventCreateWithFlags(&p->ev_main[k], hipEventDisableTiming));
    if (!p->ev_fin[k]) GL_HIP(hipEventCreateWithFlags(&p->ev_fin[k], hipEventDisableTiming));
  }
  // default priority; non-blocking so that it overlaps a caller on the
  // legacy default stream
  if (!p->side) GL_HIP(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
  p->overlap = true;
  return GL_OK;
}

gl_status gl_trace_plan_flush(gl_trace_plan* p, void* stream) {
  if (!p) return fail(GL_ERR_INVALID_ARGUMENT, "plan is required");
  return join_side(p, (hipStream_t)stream);
}

gl_status gl_trace_plan_set_persistent_config(gl_trace_plan* p, const gl_persistent_config* c) {
  if (!p || !c) return fail(GL_ERR_INVALID_ARGUMENT, "plan and config are required");
  gl_status s = check_config(*c);
  if (s) return s;
  p->pcfg = *c;
  return GL_OK;
}

gl_status gl_trace_plan_get_persistent_config(const gl_trace_plan* p, gl_persistent_config* c) {
  if (!p || !c) return fail(GL_ERR_INVALID_ARGUMENT, "plan and config are required");
  *c = p->pcfg;
  return GL_OK;
}

gl_status gl_trace_encode_work(gl_trace_plan* p, const gl_work_item* items, size_t n_items,
                               void* stream) {
  if (!p) return fail(GL_ERR_INVALID_ARGUMENT, "plan is required");
  std::vector<FrameScene> table;
  unsigned long long total = 0, batch_first = 0;
  size_t rays = 0;
  gl_status s = build_work_table(p, items, n_items, &table, &total, &rays, &batch_first);
  if (s) return s;
  hipStream_t st = (hipStream_t)stream;
  s = upload_frame(p, table, st);
  if (s) return s;
  return launch_persistent(p, p->d_frame, table.size(), total, rays, st, p->pcfg, p->ws,
                           batch_first);
}

gl_status gl_geodesic_batch_trace_persistent(gl_trace_plan* p, uint32_t scene,
                                             const double* states8, const float* directions,
                                             size_t n, const gl_persistent_config* config,
                                             const gl_ray_outputs* out, void* stream) {
  if (!p || !out) return fail(GL_ERR_INVALID_ARGUMENT, "plan and outputs are required");
  if (scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
  const gl_persistent_config c = config ? *config : gl_persistent_config{256, 2, 8, 16, 2};
  gl_status s = check_config(c);
  if (s) return s;
  if (c.refill_threshold >= 64u)
    return fail(GL_ERR_INVALID_ARGUMENT, "geodesic batches refill lanes: refill_threshold must be below 64");
  if (n == 0) return GL_OK;
  if (!states8 || !directions) return fail(GL_ERR_INVALID_ARGUMENT, "states and directions are required");
  if (n > 0xffffffffull) return fail(GL_ERR_CAPACITY, "more than 2^32 requests");
  s = check_outputs(out, n);
  if (s) return s;
  hipStream_t st = (hipStream_t)stream;
  FrameScene f = make_frame(scene, 0, uint32_t(n), *out);
  f.states = states8;
  f.dirs = directions;
  s = upload_frame(p, {f}, st);
  if (s) return s;
  return launch_persistent(p, p->d_frame, 1, n, n, st, c, p->ws, n);
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
  const bool f64 = scene_f64_horizon(p->scenes[scene]), refine = scene_refine(p->scenes[scene]);
  s = drain_side(p);
  if (s) return s;
  s = ensure_workspace(p->ws, n, f64, refine);
  if (s) return s;
  FrameScene f = make_frame(scene, 0, uint32_t(n), *out);
  f.states = states8;  // the refine pass rebuilds a request's start from its state
  f.dirs = directions;
  s = upload_frame(p, {f}, st);
  if (s) return s;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, p->d_queue, 0ull, 0ull, p->ws.counters);
  const DeferBuf df = workspace_view(p->ws, f64, refine);
  mark(p, 0, st);
  hipLaunchKernelGGL(k_geodesic_batch, dim3(uint32_t((n + 63) / 64)), dim3(64), 0, st,
                     p->d_scenes, p->d_frame, states8, directions, uint32_t(n), map_view(p), df);
  GL_HIP(hipGetLastError());
  mark(p, 1, st);
  return launch_f64_passes(p, p->d_frame, df, st);
}

// ---------------------------------------------------------------------------
// Allocation-free geodesic batch. The frame record, the four workspace
// counters and the deferral and refine records of up to kNoAllocMaxRays
// requests live in one file-scope device array, `scratch` (about 6.8 MB,
// held on each GPU the library runs on until the process ends). Its device
// address is taken once per device with hipGetSymbolAddress; host code uses
// only that address. Every call on a GPU shares the array: calls that
// overlap on one GPU race.
namespace gl {

constexpr uint32_t kNoAllocMaxRays = 32768u;
constexpr size_t align256(size_t v) { return (v + 255u) & ~size_t(255u); }
constexpr size_t kNoAllocCounters = align256(sizeof(FrameScene));
constexpr size_t kNoAllocDefer = kNoAllocCounters + align256(kCounterWords * sizeof(unsigned int));
constexpr size_t kNoAllocOrder = align256(kNoAllocDefer + kNoAllocMaxRays * sizeof(DeferRecord));
constexpr size_t kNoAllocRefine = align256(kNoAllocOrder + kNoAllocMaxRays * sizeof(uint32_t));
constexpr size_t kNoAllocRefineOrder =
    align256(kNoAllocRefine + kNoAllocMaxRays * sizeof(RefineRecord));
constexpr size_t kNoAllocBytes =
    align256(kNoAllocRefineOrder + kNoAllocMaxRays * sizeof(uint32_t));
static_assert(alignof(FrameScene) <= 256 && alignof(DeferRecord) <= 256 &&
                  alignof(RefineRecord) <= 256,
              "scratch regions are 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kNoAllocDevices] = {};

// The current device's address of `scratch`, looked up on the first call on
// that device.
gl_status scratch_address(void** out) {
  int dev = -1;
  GL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices)
    return fail(GL_ERR_CAPACITY, "device ordinal beyond the allocation-free path's table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_address[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      return fail(GL_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(e));
    g_scratch_address[dev] = ptr;
  }
  *out = g_scratch_address[dev];
  return GL_OK;
}

// Writes the launch's frame record and clears the counters (a kernel, so a
// captured call replays both).
__global__ void k_noalloc_begin(FrameScene f, FrameScene* frame, unsigned int* counters) {
  if (blockIdx.x != 0) return;
  clear_counters(counters);
  if (threadIdx.x == 0) *frame = f;
}

}  // namespace gl

gl_status gl_geodesic_batch_trace_noalloc(gl_trace_plan* p, uint32_t scene, const double* states8,
                                          const float* directions, size_t n,
                                          const gl_ray_outputs* out, void* stream) {
  if (!p || !out) return fail(GL_ERR_INVALID_ARGUMENT, "plan and outputs are required");
  if (scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
  if (n == 0) return GL_OK;
  if (!states8 || !directions) return fail(GL_ERR_INVALID_ARGUMENT, "states and directions are required");
  if (n > kNoAllocMaxRays)
    return fail(GL_ERR_CAPACITY, "the allocation-free geodesic batch takes at most 32768 requests");
  gl_status s = check_outputs(out, n);
  if (s) return s;
  void* base = nullptr;
  s = scratch_address(&base);
  if (s) return s;
  unsigned char* b = static_cast<unsigned char*>(base);
  FrameScene* frame = reinterpret_cast<FrameScene*>(b);
  unsigned int* counters = reinterpret_cast<unsigned int*>(b + kNoAllocCounters);
  const bool f64 = scene_f64_horizon(p->scenes[scene]), refine = scene_refine(p->scenes[scene]);
  DeferBuf df;
  df.rec = f64 ? reinterpret_cast<DeferRecord*>(b + kNoAllocDefer) : nullptr;
  df.order = f64 ? reinterpret_cast<uint32_t*>(b + kNoAllocOrder) : nullptr;
  df.count = counters;
  df.capacity = f64 ? kNoAllocMaxRays : 0u;
  df.refine = refine ? reinterpret_cast<RefineRecord*>(b + kNoAllocRefine) : nullptr;
  df.refine_order = refine ? reinterpret_cast<uint32_t*>(b + kNoAllocRefineOrder) : nullptr;
  df.refine_capacity = refine ? kNoAllocMaxRays : 0u;
  FrameScene f = make_frame(scene, 0, uint32_t(n), *out);
  f.states = states8;
  f.dirs = directions;
  hipStream_t st = (hipStream_t)stream;
  s = join_side(p, st);  // no allocation here: the stream waits, the host does not
  if (s) return s;
  hipLaunchKernelGGL(k_noalloc_begin, dim3(1), dim3(64), 0, st, f, frame, counters);
  GL_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_geodesic_batch, dim3(uint32_t((n + 63) / 64)), dim3(64), 0, st,
                     p->d_scenes, frame, states8, directions, uint32_t(n), map_view(p), df);
  GL_HIP(hipGetLastError());
  return launch_f64_kernels(p, frame, df, st);
}

gl_status gl_geodesic_batch_trace_f64(gl_trace_plan* p, uint32_t scene, const double* states8,
                                      const float* directions, size_t n, double rtol,
                                      double atol, const gl_ray_outputs* out, void* stream) {
  if (!p || !out) return fail(GL_ERR_INVALID_ARGUMENT, "plan and outputs are required");
  if (scene >= p->scenes.size()) return fail(GL_ERR_INVALID_ARGUMENT, "scene index out of range");
  if (!(rtol >= 1e-15 && rtol <= 1e-3) || !(atol >= 1e-18 && atol <= 1e-3))
    return fail(GL_ERR_INVALID_ARGUMENT, "f64 batch tolerances: rtol in [1e-15, 1e-3], atol in [1e-18, 1e-3]");
  if (n == 0) return GL_OK;
  if (!states8 || !directions) return fail(GL_ERR_INVALID_ARGUMENT, "states and directions are required");
  if (n > 0xffffffffull) return fail(GL_ERR_CAPACITY, "more than 2^32 requests");
  gl_status s = check_outputs(out, n);
  if (s) return s;
  hipStream_t st = (hipStream_t)stream;
  s = drain_side(p);
  if (s) return s;
  s = ensure_workspace(p->ws, 1, false, false);
  if (s) return s;
  FrameScene f = make_frame(scene, 0, uint32_t(n), *out);
  f.states = states8;
  f.dirs = directions;
  s = upload_frame(p, {f}, st);
  if (s) return s;
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, p->d_queue, 0ull, 0ull, p->ws.counters);
  mark(p, 0, st);
  hipLaunchKernelGGL(k_batch_f64, dim3(std::max(1, p->cu_count * 4)), dim3(64), 0, st,
                     p->d_scenes, p->d_frame, map_view(p), p->ws.counters + kBatchCursor,
                     uint32_t(n), rtol, atol);
  GL_HIP(hipGetLastError());
  mark(p, 1, st);
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
  s = drain_side(p);  // nothing of the second stream in flight across the capture
  if (s) return s;
  auto* g = new gl_frame_graph();
  g->plan = p;
  // The graph's own workspace, allocated before capture: eager calls on the
  // plan may later grow or replace the plan's without affecting replays.
  s = ensure_workspace(g->ws, rays, any_f64_horizon(p), any_refine(p));
  if (s) {
    gl_frame_graph_destroy(g);
    return s;
  }
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
  s = launch_persistent(p, g->d_table, table.size(), total, rays, st, p->pcfg, g->ws, total);
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
  gl_status s = join_side(g->plan, (hipStream_t)stream);
  if (s) return s;
  GL_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return GL_OK;
}

gl_status gl_frame_graph_destroy(gl_frame_graph* g) {
  if (!g) return GL_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->d_table) (void)hipFree(g->d_table);
  free_workspace(g->ws);
  delete g;
  return GL_OK;
}

}  // extern "C"
