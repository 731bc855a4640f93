/*
 * gravlume.h - C ABI of the MI355X (gfx950) Kerr-Newman null-geodesic trace.
 *
 * Drop-in boundary for the reference's two trace seams:
 *   - the GPU trace dispatch of gravlume-render
 *       TracePipeline::new      crates/gravlume-render/src/trace.rs:320-327
 *       TracePipeline::create_target                     trace.rs:565-663
 *       TracePipeline::encode_batch (+ TileRegion)       trace.rs:701-857
 *   - the CPU batch API of gravlume-reference
 *       GeodesicBatch::new / trace_ordered crates/gravlume-reference/src/batch.rs:19-43
 * Plain C: no torch or HIP types in the signatures. Device buffers are
 * caller-owned device pointers; `stream` is a hipStream_t passed as void*
 * (NULL = the default stream). Every entry point returns a gl_status; the
 * text of the last error on the calling thread is gl_last_error().
 */
#ifndef GRAVLUME_H
#define GRAVLUME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gl_status {
  GL_OK = 0,
  GL_ERR_INVALID_ARGUMENT = 1,    /* null pointer, size or enum out of range */
  GL_ERR_DOMAIN = 2,              /* scene fails a physical validity check   */
  GL_ERR_NOT_REPRESENTABLE = 3,   /* a value does not survive the f32 packing */
  GL_ERR_DEVICE = 4,              /* HIP runtime failure                     */
  GL_ERR_CAPACITY = 5             /* request exceeds a documented limit      */
} gl_status;

/* Termination classes (reference GPU protocol trace_protocol.wgsl:85-91 and
 * oracle Termination, crates/gravlume-reference/src/outcome.rs:149-158). */
enum {
  GL_TERM_ACTIVE = 0,
  GL_TERM_HORIZON = 1,
  GL_TERM_ESCAPE = 2,
  GL_TERM_SINGULARITY = 3,
  GL_TERM_STEP_EXHAUSTION = 4,
  GL_TERM_NUMERICAL_FAILURE = 5,
  GL_TERM_UNCERTAIN = 6,
  GL_TERM_SURFACE = 7,
  GL_TERM_REJECT_EXHAUSTION = 8
};

/* The status word of a ray (gl_ray_outputs.status):
 *   bits 0-3   termination class (GL_TERM_*)
 *   bits 4-7   event candidates (GL_EVENT_*)
 *   bits 8-15  numerical flags (GL_FLAG_*)
 *   bits 16-19 for GL_TERM_UNCERTAIN: the definite class the trace reached
 *              before the downgrade (0 otherwise)
 *   bit  20    GL_STATUS_NEAR_CRITICAL: the ray's distance from the critical
 *              curve (the radial-potential margin of its constants of
 *              motion) lies below the trace's own error budget
 *   bit  21    GL_STATUS_REFINED: the ray was traced again in f64 at the
 *              scene's refine tolerance (GL_OPT_NEAR_CRITICAL_REFINE), or in
 *              f64 throughout by gl_geodesic_batch_trace_f64
 *   bit  22    GL_STATUS_BRANCH_UNSETTLED: the margin lies below
 *              branch_factor times the error budget, so the branch key
 *              (turnings, crossings, winding), the travel time and the
 *              azimuth advance, which grow with the time spent near the
 *              critical orbit, are not settled; the class may still be
 *              definite (the margin lies above near_critical_factor times
 *              the budget)
 *   bit  23    GL_STATUS_NO_ADJOINT (gl_vjp_forward only): a surface ray the
 *              adjoint gives no gradient, because its trace took more
 *              accepted steps than the adjoint records or entered the f64
 *              near-horizon island, which the adjoint forward does not use
 *              (its outputs there differ from the trace's)
 * A ray is GL_TERM_UNCERTAIN when its terminal events tie
 * (GL_FLAG_TIED_EVENTS), a radial turning lies inside the tie window of the
 * terminal (GL_FLAG_TURNING_ORDER), an invariant drift exceeds the scene
 * budget (GL_FLAG_DRIFT_BUDGET) (geodesic_integration.wgsl:324-330), or it is
 * near-critical and not refined (docs/validation.md:287). */
#define GL_STATUS_PROVISIONAL_SHIFT 16
#define GL_STATUS_NEAR_CRITICAL (1u << 20)
#define GL_STATUS_REFINED (1u << 21)
#define GL_STATUS_BRANCH_UNSETTLED (1u << 22)
#define GL_STATUS_NO_ADJOINT (1u << 23)

/* Event-candidate bits (trace_protocol.wgsl:95-98, geodesic.rs:616-634). */
enum {
  GL_EVENT_SINGULARITY = 1u,
  GL_EVENT_HORIZON = 2u,
  GL_EVENT_SURFACE = 4u,
  GL_EVENT_ESCAPE = 8u
};

/* Numerical flags (trace_protocol.wgsl:92-94). */
enum {
  GL_FLAG_NON_FINITE = 1u,
  GL_FLAG_INVALID_RADICAND = 2u,
  GL_FLAG_INVALID_DENOMINATOR = 4u,
  GL_FLAG_DRIFT_BUDGET = 8u,     /* invariant drift above the scene limit    */
  GL_FLAG_TURNING_ORDER = 16u,   /* turning bracket overlaps the terminal    */
  GL_FLAG_TIED_EVENTS = 32u,     /* several event candidates within the tie  */
  GL_FLAG_CERTIFIED_CAPTURE = 64u, /* horizon class proven by the interval
                                      capture certificate, not integrated  */
  GL_FLAG_RECONSTRUCTED = 128u      /* escape direction interpolated from the
                                      escape-direction stencil, not traced:
                                      steps, times, drifts and branch key 0 */
};

/* Option bits of gl_scene_desc.options. */
enum {
  GL_OPT_CAPTURE_CERTIFICATE = 1u, /* interval radial-capture certificate
                                      (scenes without a surface): proven
                                      captures end at once as class HORIZON
                                      with GL_FLAG_CERTIFIED_CAPTURE */
  GL_OPT_F64_HORIZON_STEPS = 2u,   /* f64 DP5 steps inside 1.05 r_+            */
  GL_OPT_F64_DISK_ROOT = 4u,       /* f64 equatorial root solve + observables  */
  GL_OPT_ESCAPE_MAP = 8u,          /* look escape directions up in the map     */
  GL_OPT_NEAR_CRITICAL_REFINE = 16u /* near-critical rays (bit 20 of the status)
                                      are traced again from f64 initial data in
                                      f64 at refine_rtol / refine_atol; they
                                      stay GL_TERM_UNCERTAIN only if the refined
                                      trace is still near-critical */
};

/* Surface emission and transport models (EquatorialEmissionModel,
 * emission.rs:3-10; SurfaceTransport, scene.rs:40-43). */
enum { GL_EMISSION_BOLOMETRIC = 0u, GL_EMISSION_BLACKBODY = 1u };
enum { GL_TRANSPORT_VACUUM = 0u, GL_TRANSPORT_SLAB = 1u };

/* The blackbody band table (spectral_lut.rs:5-10): entry i holds the log2
 * fractions (red 600-700 nm, green 500-600 nm, blue 400-500 nm, 0) of a
 * blackbody at T = 2^(GL_BLACKBODY_LUT_MIN_LOG2_K + i / PER_OCTAVE) K. */
#define GL_BLACKBODY_LUT_ENTRIES 4097
#define GL_BLACKBODY_LUT_MIN_LOG2_K (-8.0)
#define GL_BLACKBODY_LUT_PER_OCTAVE 128.0

/* One scene: the content of the reference's validated Observation
 * (crates/gravlume-domain/src/scene.rs:99-143, view.rs:28-69), its event
 * configuration (crates/gravlume-reference/src/events.rs:17-105) and the
 * integrator policy. All lengths in units of M; mass must be exactly 1
 * (geodesic.rs:28-41). */
typedef struct gl_scene_desc {
  double mass, spin, charge;
  int32_t chart; /* +1 ingoing, -1 outgoing Kerr-Schild chart */
  uint32_t options;
  double observer_txyz[4]; /* stationary observer event (Cartesian KS) */
  double target_txyz[4];   /* image-centre target */
  double up_hint[3];       /* image-up hint */
  double frequency;        /* observer frequency, must be 1 (input.rs:57-63) */
  uint32_t width, height;  /* viewport; pixel (0,0) is top-left; sides <= 65536,
                              width * height <= 2^32 - 1 (else GL_ERR_CAPACITY) */
  double vertical_fov;     /* radians, in (0, pi) */
  double subpixel[2];      /* sample offset inside each pixel, in [0,1] */
  double escape_radius;    /* escape event r = escape_radius */
  double horizon_offset;   /* capture event r = r_+ (1 + horizon_offset) */
  int32_t has_surface;     /* equatorial prograde Keplerian emitter */
  double surface_inner, surface_outer;
  /* Emitter (crates/gravlume-domain/src/emission.rs:12-145): bolometric
   * intensity I(r) = intensity_at_six (r / 6)^-3; with GL_EMISSION_BLACKBODY
   * a diluted blackbody at T(r) = temperature_at_six (r / 6)^-3/4 [K]. */
  uint32_t emission_model;  /* GL_EMISSION_BOLOMETRIC or GL_EMISSION_BLACKBODY */
  uint32_t transport_model; /* GL_TRANSPORT_VACUUM or GL_TRANSPORT_SLAB */
  double intensity_at_six;
  double temperature_at_six;
  /* Homogeneous slab between source and observer (HomogeneousScalarSlab,
   * crates/gravlume-domain/src/radiation.rs:54-220):
   * I_out = I_in exp(-optical_depth) + slab_emission, the emission term a
   * blackbody at slab_temperature [K] (0: neutral bolometric). */
  double optical_depth;
  double slab_emission;
  double slab_temperature;
  /* adaptive Dormand-Prince 5(4) policy */
  double rtol, atol;
  double initial_step, minimum_step, maximum_step;
  double maximum_step_per_radius; /* |h| <= this * r (0 disables) */
  /* step controller h <- h clamp(safety E^-1/5, min_factor, max_factor)
   * (ReferencePolicy, policy.rs:5-66, 160-173); 0 selects the reference's
   * 0.9, 0.2 and 5 */
  double step_safety, step_min_factor, step_max_factor;
  uint32_t max_accepted_steps, max_consecutive_rejects;
  double event_tie_tolerance; /* M */
  double arming_band;         /* M */
  double drift_limit;         /* relative invariant drift budget */
  uint32_t reproject_interval; /* f64 null re-projection every N DP5 attempts
                                  (accepted or rejected) of a ray; 0 off */
  uint32_t reserved;
  /* Near-critical test: a ray that came within 1.25 r_m of the radius r_m
   * where its radial potential R(r) is smallest over [r_+, r_start] is
   * near-critical when |R(r_m)| / (sum of the magnitudes of R's two terms at
   * r_m) <= near_critical_factor * (the trace's error budget: the largest of
   * its relative null, L_z and Carter drifts sampled at r from
   * max(7/8 r_m, (r_m + r_+) / 2) to 2 r_m,
   * and for an f32 trace the f32 rounding of its initial constants). 0
   * selects 16; negative disables the test (and the branch test). Scenes
   * without a horizon are never near-critical. */
  double near_critical_factor;
  /* the same test for the branch key and the orbit-dependent observables
   * (GL_STATUS_BRANCH_UNSETTLED); 0 selects 4096, must be >= the
   * near-critical factor */
  double branch_factor;
  /* f64 refine policy of GL_OPT_NEAR_CRITICAL_REFINE (which re-traces every
   * ray with bit 20 or 22); 0 selects 1e-12 / 1e-14 */
  double refine_rtol, refine_atol;
} gl_scene_desc;

/* Rectangle of 8x8 tiles of one scene (TileRegion, trace.rs:799-857). */
typedef struct gl_tile_region {
  uint32_t scene;
  uint32_t tile_x, tile_y;
  uint32_t tiles_x, tiles_y;
} gl_tile_region;

/* Structure-of-arrays per-ray outputs, caller-owned device memory, one
 * element per ray (index = y * width + x for image traces, request order
 * for geodesic batches). Any pointer may be NULL to skip that plane.
 *   status      bits 0-3 termination, 4-7 event candidates, 8-15 flags
 *   steps       accepted DP5 steps; rejected in `rejected`
 *   observable  escape: unit traversal direction (x,y,z);
 *               surface: (r, phi_s, g) source anchor and frequency ratio
 *   travel_time coordinate-time duration along the traversal (>= 0)
 *   affine      affine parameter at termination (negative for backward)
 *   drift       (max null residual, max |dL_z|, max |dQ|) relative drifts
 *   branch      bits 0-7 radial turnings, 8-15 equatorial crossings,
 *               16-23 azimuth winding (two's complement), 24-25 polar side
 *   radiance    escape: escape-map RGB (0 without a map); surface: the
 *               observed radiance (surface_transport.wgsl:209-223,
 *               blackbody_surface_preview.wgsl:37-80): red, green, blue
 *               band intensities of a blackbody emitter, or the neutral
 *               bolometric intensity in all three; -1 when the radiance is
 *               not representable in RGBA16F; 0 for every other class
 *   rgba16f     the scene-linear RGBA16F texel, 4 halves per ray in R, G, B,
 *               A order (lensing_preview.wgsl:64-82): surface (radiance, 2),
 *               escape (map RGB, 1), horizon (0, 0, 0, 0), anything else a
 *               failure colour with alpha -termination
 *   event_residual  the localized event function re-evaluated at the
 *               terminal state, (value - event surface) / max(1, |surface|):
 *               r - r_escape or r - r_capture, z on the disk, the singularity
 *               measure minus its guard; 0 for a terminal without an event
 *               (GeometricSample.event_residual, trace_protocol.wgsl:76)
 *   turning_radius  the smallest radius at a localized inward-to-outward
 *               radial turning, -1 when the ray made none
 *               (ReferenceOutcome::turning_radius_m, outcome.rs:425-436)
 *   azimuth_advance the signed azimuth swept from the initial to the terminal
 *               state, the sum of wrapped atan2(y, x) differences between
 *               committed states (ReferenceOutcome::azimuth_advance_rad) */
typedef struct gl_ray_outputs {
  uint32_t* status;
  uint32_t* steps;
  uint32_t* rejected;
  float* observable; /* 3 planes of n floats: [0,n) [n,2n) [2n,3n) */
  float* travel_time;
  float* affine;
  float* drift;      /* 3 planes */
  uint32_t* branch;
  float* radiance;   /* 3 planes */
  uint16_t* rgba16f; /* 4 halves per ray, interleaved */
  float* event_residual;
  float* turning_radius;  /* the turning localization runs only when given */
  float* azimuth_advance;
  size_t count;      /* number of rays the planes hold */
} gl_ray_outputs;

typedef struct gl_trace_plan gl_trace_plan;

/* Host-side spectral helpers (no device needed). gl_blackbody_log2_fractions
 * gives log2 of the fraction of sigma T^4 in each visible band at T [K]
 * (blackbody_band_log2_fractions, spectral_lut.rs:72-88);
 * gl_blackbody_lut fills the GL_BLACKBODY_LUT_ENTRIES x 4 floats the kernels
 * interpolate (blackbody_log2_fraction_lut, spectral_lut.rs:40-56). */
gl_status gl_blackbody_log2_fractions(double temperature_kelvin, double* out_log2);
gl_status gl_blackbody_lut(float* out, size_t n_floats);

/* The scene checks of gl_trace_plan_create for one scene, on the host only
 * (no device needed): GL_OK or the status and gl_last_error() text the plan
 * would fail with (input.rs:45-131, observer.rs:67-183). */
gl_status gl_scene_validate(const gl_scene_desc* scene);

/* Validates and packs n_scenes scenes (TracePipeline::new, input.rs:45-131)
 * and uploads them with the escape map (width x height texels, 3 float
 * planes on the host, may be NULL). */
gl_status gl_trace_plan_create(const gl_scene_desc* scenes, size_t n_scenes,
                               const float* escape_map_rgb, uint32_t map_width,
                               uint32_t map_height, gl_trace_plan** out_plan);
gl_status gl_trace_plan_destroy(gl_trace_plan* plan);

/* Number of rays of a scene (width * height). */
gl_status gl_trace_plan_scene_rays(const gl_trace_plan* plan, uint32_t scene, size_t* out_rays);

/* Plain trace: one ray per lane over the tiles of `region`, 8x8 tiles per
 * 64-lane wavefront (encode_batch, trace.rs:701-784). */
gl_status gl_trace_encode(gl_trace_plan* plan, const gl_tile_region* region,
                          const gl_ray_outputs* out, void* stream);

/* Persistent trace: every ray of the listed scenes (all scenes if scenes is
 * NULL) in one launch from a device work queue with wavefront refill.
 * outputs[i] receives scene scenes[i]. */
gl_status gl_trace_encode_persistent(gl_trace_plan* plan, const uint32_t* scenes,
                                     size_t n_scenes, const gl_ray_outputs* outputs,
                                     void* stream);

/* Mixed work list for one persistent launch: whole image scenes and
 * geodesic batches of any scenes of the plan, all drawn from the same work
 * queue (uneven multi-scene batches in one launch). In whole-tile mode
 * (refill_threshold 64) with more than one item, the image scenes' 8x8 tiles
 * are handed out whole first and the batch requests after them one lane at a
 * time (lane refill in rounds of 16 attempts at 16 free lanes): requests of
 * one batch differ in length by orders of magnitude, pixels of one tile do
 * not. Each ray's outputs are the same in every mode. */
enum { GL_WORK_IMAGE = 0u, GL_WORK_GEODESIC_BATCH = 1u };
typedef struct gl_work_item {
  uint32_t scene;
  uint32_t kind;            /* GL_WORK_IMAGE or GL_WORK_GEODESIC_BATCH */
  const double* states8;    /* batch: n x (t, x, y, z, p_t, p_x, p_y, p_z), device */
  const float* directions;  /* batch: n affine directions, device */
  size_t count;             /* batch: n requests (ignored for images) */
  gl_ray_outputs outputs;   /* image: width*height rays; batch: n, request order */
} gl_work_item;
gl_status gl_trace_encode_work(gl_trace_plan* plan, const gl_work_item* items, size_t n_items,
                               void* stream);

/* Accelerated sky trace (trace_scene_accelerated and trace_escape_map_nodes,
 * geodesic_acceleration.wgsl:330-601) of one scene without a surface: the
 * pixels (4i, 4j) are traced in full, then every 8x8 tile whose 3x3 such
 * nodes all escape and agree with the bilinear blend of its corners within
 * min(half a pixel, 3e-4 rad) gets the blended escape direction for its
 * other pixels (GL_FLAG_RECONSTRUCTED); all other pixels are traced in full.
 * Needs the observable plane. */
gl_status gl_trace_encode_accelerated(gl_trace_plan* plan, uint32_t scene,
                                      const gl_ray_outputs* out, void* stream);

/* Shadow-edge coverage (classify_shadow_edges and refine_shadow_edges,
 * shadow_coverage.wgsl:1-136) on a traced image of `scene`: every pixel of
 * class horizon or escape with a 4-neighbour of the other class is re-traced
 * at four rotated-grid sub-samples (f32); its RGBA16F texel becomes the mean
 * escape colour with alpha = escaping fraction, unless a sub-sample ends in
 * another class. Reads the status plane, rewrites RGBA16F texels only. */
gl_status gl_trace_refine_shadow(gl_trace_plan* plan, uint32_t scene, const gl_ray_outputs* out,
                                 void* stream);
/* Edge pixels listed by the last gl_trace_refine_shadow (synchronizes the stream). */
gl_status gl_trace_plan_shadow_edges(gl_trace_plan* plan, uint64_t* count, void* stream);

/* Source-chart footprint of surface pixels (surface_footprint_v1,
 * crates/gravlume-reference/src/footprint.rs:74-151; GPU capture
 * surface_footprint_capture.wgsl:1-99): for each pixel (the device list
 * `pixels`, indices y * width + x, or every pixel when NULL and n equals the
 * pixel count) the centre and its four +-0.25-pixel neighbours are traced;
 * when all five end unambiguously on the surface (no flags, candidate mask
 * exactly the surface bit) with one branch key, jacobian[4i..4i+3] =
 * (dr/dx, dr/dy, d(r_c dphi)/dx, d(r_c dphi)/dy) in M per pixel by central
 * differences, else zeros. meta[i]: bit 0 resolved, bits 1-2 parity (1
 * positive, 2 negative, 3 degenerate), bits 8-11 the centre's termination,
 * bits 16-23 its flags. An index outside the image is not traced (zeros,
 * meta 0). Device outputs. */
gl_status gl_trace_surface_footprint(gl_trace_plan* plan, uint32_t scene, const uint32_t* pixels,
                                     size_t n, float* jacobian, uint32_t* meta, void* stream);

/* Launch parameters of the persistent kernel (set by the autotuner; a new
 * plan starts in whole-tile mode, {256, 2, 8, 64, 2}). */
typedef struct gl_persistent_config {
  uint32_t block_size;       /* lanes per workgroup: 64, 128, 192 or 256 */
  uint32_t blocks_per_cu;    /* resident workgroups per CU (capped by occupancy) */
  uint32_t refill_interval;  /* DP5 attempts between refill checks, rounded
                                up to a multiple of 16 (so the rays of a
                                wavefront keep their drift and re-projection
                                clocks in phase) */
  uint32_t refill_threshold; /* free lanes that trigger a lane refill, 1..63;
                                64 = whole-tile mode: a wavefront takes one
                                8x8 tile at a time and finishes it first */
  uint32_t waves_per_simd;   /* register budget: 2, 3 or 4 resident waves per
                                SIMD (2 or 3 in whole-tile mode) */
} gl_persistent_config;
gl_status gl_trace_plan_set_persistent_config(gl_trace_plan* plan,
                                              const gl_persistent_config* config);
gl_status gl_trace_plan_get_persistent_config(const gl_trace_plan* plan,
                                              gl_persistent_config* config);

/* f64 overlap for frame sequences (off by default; no reference counterpart:
 * the reference renders one frame at a time). When enabled, an eager
 * persistent launch (gl_trace_encode_persistent, gl_trace_encode_work,
 * gl_geodesic_batch_trace_persistent) of a plan
 * whose scenes defer near-horizon steps or refine near-critical rays runs its
 * f64 passes (near-horizon finisher, refine) on a second stream the plan
 * owns, beside the next launch's trace. A launch's outputs are then complete
 * on the caller's stream only after the NEXT launch's trace kernel, or after
 * gl_trace_plan_flush on that stream; every other entry point, and disabling
 * the mode, joins the outstanding passes first. Until then the launch's
 * inputs and output planes must stay untouched, so two consecutive
 * overlapped launches write different output planes (alternate two
 * targets). Results are bit-identical to the serial launches. */
gl_status gl_trace_plan_set_f64_overlap(gl_trace_plan* plan, int enable);
/* Makes `stream` wait for every f64 pass still running on the plan's second
 * stream (no host wait). */
gl_status gl_trace_plan_flush(gl_trace_plan* plan, void* stream);

/* Geodesic batch (GeodesicBatch::trace_ordered, batch.rs:31-43) on the GPU:
 * n requests of scene `scene`'s spacetime, policy and events, each an
 * initial state (t, x, y, z, p_t, p_x, p_y, p_z) in f64 on the device and an
 * affine direction (+1 forward, -1 backward). Results in request order. */
gl_status gl_geodesic_batch_trace(gl_trace_plan* plan, uint32_t scene,
                                  const double* states8, const float* directions, size_t n,
                                  const gl_ray_outputs* out, void* stream);

/* The same batch through the persistent work queue: free lanes refill with
 * new requests (lane-refill mode; config->refill_threshold < 64, NULL =
 * {256, 2, 8, 16, 2}), so requests of very different lengths (near-critical
 * rays) share wavefronts without waiting for the longest of 64. Results are
 * identical in meaning to gl_geodesic_batch_trace. No reference counterpart
 * beyond GeodesicBatch::trace_ordered (batch.rs:31-43). */
gl_status gl_geodesic_batch_trace_persistent(gl_trace_plan* plan, uint32_t scene,
                                             const double* states8, const float* directions,
                                             size_t n, const gl_persistent_config* config,
                                             const gl_ray_outputs* out, void* stream);

/* The geodesic batch in f64 throughout (GeodesicBatch::trace_ordered,
 * batch.rs:31-43, with the reference tracer's arithmetic, geodesic.rs:129-212):
 * every request is traced by f64 DP5(4) from its exact initial state at the
 * given tolerances (rtol in [1e-15, 1e-3], atol in [1e-18, 1e-3]; the scene
 * supplies the step policy and the events), with the near-critical test of
 * an exact start and f64 drifts. Results in request order, marked
 * GL_STATUS_REFINED. Device inputs and outputs as for
 * gl_geodesic_batch_trace. */
gl_status gl_geodesic_batch_trace_f64(gl_trace_plan* plan, uint32_t scene,
                                      const double* states8, const float* directions, size_t n,
                                      double rtol, double atol, const gl_ray_outputs* out,
                                      void* stream);

/* gl_geodesic_batch_trace without any device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. The
 * frame record, counters and f64 deferral/refine records live in a static
 * device array of about 6.8 MB per GPU (held until the process ends); the
 * caller passes the plan, inputs and outputs as for gl_geodesic_batch_trace,
 * and gets the same results. At most 32768 requests per call (else
 * GL_ERR_CAPACITY). Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads) race. */
gl_status gl_geodesic_batch_trace_noalloc(gl_trace_plan* plan, uint32_t scene,
                                          const double* states8, const float* directions,
                                          size_t n, const gl_ray_outputs* out, void* stream);

/* Persistent-queue statistics of the last persistent launch. */
typedef struct gl_queue_stats {
  uint64_t items;          /* rays handed out */
  uint64_t refills;        /* wavefront refill rounds */
  uint64_t attempted_steps;/* DP5 attempts over all lanes */
  uint64_t lane_slots;     /* lane-attempt slots offered (active or idle) */
} gl_queue_stats;
gl_status gl_trace_plan_queue_stats(gl_trace_plan* plan, gl_queue_stats* stats, void* stream);

/* Kernel timing: when enabled, HIP events on the launch stream bracket the
 * main trace kernel and the f64 near-horizon finisher of every subsequent
 * encode call (not inside graph capture). gl_trace_plan_kernel_ms
 * synchronizes on those events and returns the mean durations over every
 * launch since timing was enabled or the durations were last read (then
 * starts a new count); `launches` (may be NULL) receives how many main
 * kernels the mean covers. Enabling timing clears the count. */
gl_status gl_trace_plan_enable_timing(gl_trace_plan* plan, int enable);
gl_status gl_trace_plan_kernel_ms(gl_trace_plan* plan, double* main_ms, double* finish_ms,
                                  uint64_t* launches);

/* Frame capture: records the persistent frame (queue reset + launch) into a
 * hipGraph once; gl_frame_graph_launch replays it. */
typedef struct gl_frame_graph gl_frame_graph;
gl_status gl_frame_graph_create(gl_trace_plan* plan, const uint32_t* scenes, size_t n_scenes,
                                const gl_ray_outputs* outputs, void* stream,
                                gl_frame_graph** out_graph);
gl_status gl_frame_graph_launch(gl_frame_graph* graph, void* stream);
gl_status gl_frame_graph_destroy(gl_frame_graph* graph);

/* Discrete-adjoint vector-Jacobian product of per-pixel frequency ratios g
 * with respect to (spin a, charge Q, inclination, pixel direction), following
 * the torch.autograd.Function contract: forward saves state, backward takes
 * dL/dg per pixel. Scene `scene` must be an image scene with a surface;
 * `inclination` is the observer's polar angle in radians (checked against the
 * scene's observer position, whose oblate radius and azimuth stay fixed).
 * The forward writes the trace outputs to `out` and records the accepted
 * step sizes (up to min(max_accepted_steps, 1024) per ray; longer rays get
 * no gradient), a checkpoint every `checkpoint_interval` accepted steps and
 * the terminal state; no f64 near-horizon deferral is used (surface rays
 * that entered the island get no gradient either; both carry
 * GL_STATUS_NO_ADJOINT in `out`'s status plane). The backward
 * synchronizes `stream`. With plan timing enabled, both kernels are timed
 * like a trace launch (gl_trace_plan_kernel_ms). */
typedef struct gl_vjp_state gl_vjp_state;
gl_status gl_vjp_forward(gl_trace_plan* plan, uint32_t scene, double inclination,
                         uint32_t checkpoint_interval, const gl_ray_outputs* out, void* stream,
                         gl_vjp_state** out_state);
/* grad_g: device float[n_rays]; grad_params: host double[3] (a, Q, incl);
 * grad_direction: device float[2 * n_rays] (d/d sight-plane x, y), may be NULL. */
gl_status gl_vjp_backward(gl_vjp_state* state, const float* grad_g, double* grad_params,
                          float* grad_direction, void* stream);
/* The same for a seed on any of the surface observables: grad_r, grad_phi,
 * grad_g are device float[n_rays] of dL/dr, dL/dphi_s, dL/dg (NULL = 0, at
 * least one given). Seeding r and then phi_s gives every surface pixel's
 * lensing Jacobian d(r, phi_s)/d(sight-plane x, y) in two passes. */
gl_status gl_vjp_backward_observables(gl_vjp_state* state, const float* grad_r,
                                      const float* grad_phi, const float* grad_g,
                                      double* grad_params, float* grad_direction, void* stream);
/* The source-chart footprint of every pixel of the state's scene from two
 * backward passes (seed dL/dr = 1, then dL/dphi_s = 1): the analytic form of
 * gl_trace_surface_footprint (surface_footprint_v1,
 * crates/gravlume-reference/src/footprint.rs:74-151) without its four
 * neighbour rays. `out` holds the forward's planes (status and observable
 * are read). jacobian: device float[4 * n_rays], (dr/dx, dr/dy,
 * r dphi/dx, r dphi/dy) in M per pixel (x right, y down), zeros unless
 * resolved; meta: device uint32[n_rays] as in gl_trace_surface_footprint,
 * resolved when the ray ended unambiguously on the surface (no flags,
 * candidate mask exactly the surface bit) within the adjoint's step
 * capacity. Enqueued on `stream`, no synchronization. */
gl_status gl_vjp_surface_footprint(gl_vjp_state* state, const gl_ray_outputs* out,
                                   float* jacobian, uint32_t* meta, void* stream);
gl_status gl_vjp_destroy(gl_vjp_state* state);

/* Per-thread text of the last error. */
const char* gl_last_error(void);

/* Library build identity and the counted per-RHS / per-attempt flop
 * constants of the f32 trace (used for roofline accounting). */
const char* gl_version(void);
gl_status gl_flop_constants(double* rhs_flops, double* attempt_flops);

#ifdef __cplusplus
}
#endif

#endif /* GRAVLUME_H */
