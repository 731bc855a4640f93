"""Host mirror of the reference's GPU trace seam on top of the C ABI.

Names follow crates/gravlume-render/src/trace.rs: a ``TracePipeline`` owns the
packed scenes (TracePipeline::new, trace.rs:320-327), ``create_target``
allocates the per-ray output planes (trace.rs:565-663) and ``encode_batch``
traces a ``TileRegion`` of 8x8 tiles (trace.rs:701-857). ``encode_persistent``
traces whole scenes in one persistent launch; ``GeodesicBatch`` mirrors
gravlume-reference's GeodesicBatch::trace_ordered (batch.rs:19-43) on the GPU.
PyTorch only provides device memory and the stream.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi
from .scene import Scene, escape_map


class GpuTraceInputError(ValueError):
    """The scene failed validation or the f32 packing (input.rs:342-381)."""


@dataclass
class TileRegion:
    """A rectangle of 8x8 tiles of one scene (trace.rs:799-857)."""
    scene: int
    tile_x: int
    tile_y: int
    tiles_x: int
    tiles_y: int

    @staticmethod
    def full(scene_index: int, width: int, height: int) -> "TileRegion":
        return TileRegion(scene_index, 0, 0, (width + 7) // 8, (height + 7) // 8)


class TraceTarget:
    """Caller-owned structure-of-arrays outputs for n rays on the device."""

    def __init__(self, rays: int, device: str | torch.device = "cuda", scene: int | None = None,
                 turning_radius: bool = False):
        dev = torch.device(device)
        self.rays = rays
        self.scene = scene  # the image scene it was created for (None: a batch)
        self.status = torch.zeros(rays, dtype=torch.int32, device=dev)
        self.steps = torch.zeros(rays, dtype=torch.int32, device=dev)
        self.rejected = torch.zeros(rays, dtype=torch.int32, device=dev)
        self.observable = torch.zeros(3 * rays, dtype=torch.float32, device=dev)
        self.travel_time = torch.zeros(rays, dtype=torch.float32, device=dev)
        self.affine = torch.zeros(rays, dtype=torch.float32, device=dev)
        self.drift = torch.zeros(3 * rays, dtype=torch.float32, device=dev)
        self.branch = torch.zeros(rays, dtype=torch.int32, device=dev)
        self.radiance = torch.zeros(3 * rays, dtype=torch.float32, device=dev)
        # scene-linear RGBA16F texels, 4 halves per ray (R, G, B, A)
        self.rgba16f = torch.zeros(4 * rays, dtype=torch.int16, device=dev)
        self.event_residual = torch.zeros(rays, dtype=torch.float32, device=dev)
        self.azimuth_advance = torch.zeros(rays, dtype=torch.float32, device=dev)
        # the turning localization runs only for a target that holds this plane
        self.turning_radius = (torch.zeros(rays, dtype=torch.float32, device=dev)
                               if turning_radius else None)

    PLANES = ("status", "steps", "rejected", "observable", "travel_time", "affine", "drift",
              "branch", "radiance", "rgba16f", "event_residual", "turning_radius",
              "azimuth_advance")

    def c_struct(self) -> _capi.RayOutputsC:
        o = _capi.RayOutputsC()
        for name in self.PLANES:
            t = getattr(self, name)
            setattr(o, name, t.data_ptr() if t is not None else None)
        o.count = self.rays
        return o

    def to_host(self) -> dict:
        """Decoded numpy copies of every plane."""
        n = self.rays
        st = self.status.cpu().numpy().view(np.uint32)
        br = self.branch.cpu().numpy().view(np.uint32)
        out = {
            "termination": (st & 15).astype(np.int32),
            "candidates": ((st >> 4) & 15).astype(np.int32),
            "flags": ((st >> 8) & 255).astype(np.int32),
            # for Uncertain rays the class the trace reached before the downgrade
            "provisional": ((st >> _capi.STATUS_PROVISIONAL_SHIFT) & 15).astype(np.int32),
            "near_critical": (st & _capi.STATUS_NEAR_CRITICAL) != 0,
            "refined": (st & _capi.STATUS_REFINED) != 0,
            "branch_unsettled": (st & _capi.STATUS_BRANCH_UNSETTLED) != 0,
            "no_adjoint": (st & _capi.STATUS_NO_ADJOINT) != 0,
            "event_residual": self.event_residual.cpu().numpy(),
            "azimuth_advance": self.azimuth_advance.cpu().numpy(),
            "steps": self.steps.cpu().numpy().view(np.uint32).astype(np.int64),
            "rejected": self.rejected.cpu().numpy().view(np.uint32).astype(np.int64),
            "observable": self.observable.cpu().numpy().reshape(3, n).T.copy(),
            "travel_time": self.travel_time.cpu().numpy(),
            "affine": self.affine.cpu().numpy(),
            "drift": self.drift.cpu().numpy().reshape(3, n).T.copy(),
            "radial_turnings": (br & 255).astype(np.int32),
            "equatorial_crossings": ((br >> 8) & 255).astype(np.int32),
            "azimuth_winding": ((br >> 16) & 255).astype(np.uint8).view(np.int8).astype(np.int32),
            "polar_side": ((br >> 24) & 3).astype(np.int32),
            "radiance": self.radiance.cpu().numpy().reshape(3, n).T.copy(),
            "rgba16f": self.rgba16f.cpu().numpy().view(np.float16).reshape(n, 4).copy(),
        }
        if self.turning_radius is not None:
            out["turning_radius"] = self.turning_radius.cpu().numpy()
        # the definite class of every ray: its own class, or for an Uncertain
        # ray the class it reached before the downgrade
        out["class_reached"] = np.where(out["termination"] == _capi.TERM_UNCERTAIN,
                                        out["provisional"], out["termination"])
        return out


def _stream_handle(stream) -> C.c_void_p:
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


def _footprint_result(jac: torch.Tensor, meta: torch.Tensor, n: int) -> dict:
    """Host view of a footprint launch: resolved, parity (1 positive,
    2 negative, 3 degenerate), J (n, 2, 2) with rows (radial, local arc) and
    columns (x, y), singular values (n, 2) (major; minor = |det| / major),
    the centre's termination and flags."""
    J = jac.cpu().numpy()[:4 * n].astype(np.float64).reshape(n, 2, 2)
    m = meta.cpu().numpy()[:n].view(np.uint32)
    det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
    norm2 = np.sum(J * J, axis=(1, 2))
    major = np.sqrt(0.5 * (norm2 + np.sqrt(np.maximum(norm2 * norm2 - 4 * det * det, 0.0))))
    with np.errstate(divide="ignore", invalid="ignore"):
        minor = np.where(major == 0.0, 0.0, np.abs(det) / major)
    return dict(resolved=(m & 1).astype(bool), parity=((m >> 1) & 3).astype(np.int32), J=J,
                singular=np.stack([major, minor], axis=1),
                termination=((m >> 8) & 15).astype(np.int32),
                flags=((m >> 16) & 255).astype(np.int32))


class TracePipeline:
    """Validated, packed scenes resident on the GPU (TracePipeline::new)."""

    def __init__(self, scenes: list[Scene] | Scene, escape_map_rgb: np.ndarray | None = None,
                 map_seed: int = 20261203):
        self._lib = _capi.load()
        if isinstance(scenes, Scene):
            scenes = [scenes]
        self.scenes = list(scenes)
        descs = (_capi.SceneDesc * len(self.scenes))(*[s.to_desc() for s in self.scenes])
        if escape_map_rgb is None and any(s.escape_map for s in self.scenes):
            escape_map_rgb = escape_map(map_seed)
        self.escape_map_rgb = escape_map_rgb
        mp, mw, mh = None, 0, 0
        if escape_map_rgb is not None:
            arr = np.ascontiguousarray(escape_map_rgb, dtype=np.float32)
            mp, mh, mw = arr.ctypes.data_as(C.c_void_p), arr.shape[1], arr.shape[2]
        handle = C.c_void_p()
        torch.cuda.init()
        st = self._lib.gl_trace_plan_create(descs, len(self.scenes), mp, mw, mh, C.byref(handle))
        if st != _capi.GL_OK:
            msg = self._lib.gl_last_error().decode()
            if st in (1, 2, 3, 5):
                raise GpuTraceInputError(msg)
            raise _capi.GravlumeError(st, msg)
        self._handle = handle

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._lib.gl_trace_plan_destroy(h)
            self._handle = None

    @property
    def handle(self) -> C.c_void_p:
        return self._handle

    def scene_rays(self, scene: int) -> int:
        n = C.c_size_t()
        _capi.check(self._lib.gl_trace_plan_scene_rays(self._handle, scene, C.byref(n)))
        return n.value

    def create_target(self, scene: int = 0, device="cuda") -> TraceTarget:
        return TraceTarget(self.scene_rays(scene), device, scene=scene)

    def encode_batch(self, target: TraceTarget, region: TileRegion | None = None,
                     stream=None) -> None:
        """Plain one-ray-per-lane trace of a tile region (encode_batch); by
        default every tile of the scene the target was created for."""
        if region is None:
            k = target.scene if target.scene is not None else 0
            s = self.scenes[k]
            region = TileRegion.full(k, s.width, s.height)
        r = _capi.TileRegionC(region.scene, region.tile_x, region.tile_y, region.tiles_x,
                              region.tiles_y)
        out = target.c_struct()
        _capi.check(self._lib.gl_trace_encode(self._handle, C.byref(r), C.byref(out),
                                              _stream_handle(stream)))

    def encode_persistent(self, targets: list[TraceTarget] | TraceTarget,
                          scenes: list[int] | None = None, stream=None) -> None:
        """All rays of `scenes` (default: every scene) in one persistent launch."""
        if isinstance(targets, TraceTarget):
            targets = [targets]
        outs = (_capi.RayOutputsC * len(targets))(*[t.c_struct() for t in targets])
        if scenes is None:
            ids, n = None, len(self.scenes)
        else:
            arr = (C.c_uint32 * len(scenes))(*scenes)
            ids, n = C.cast(arr, C.c_void_p), len(scenes)
        if len(targets) != n:
            raise ValueError("one target per traced scene is required")
        _capi.check(self._lib.gl_trace_encode_persistent(self._handle, ids, n, outs,
                                                         _stream_handle(stream)))

    def encode_accelerated(self, target: TraceTarget, stream=None) -> None:
        """Accelerated sky trace of the target's scene (no surface): the
        escape-direction stencil fills stable 8x8 tiles from their traced
        4-pixel nodes, the rest is traced in full
        (geodesic_acceleration.wgsl:330-601)."""
        k = target.scene if target.scene is not None else 0
        out = target.c_struct()
        _capi.check(self._lib.gl_trace_encode_accelerated(self._handle, k, C.byref(out),
                                                          _stream_handle(stream)))

    def refine_shadow(self, target: TraceTarget, stream=None) -> None:
        """Four-sample coverage of the traced image's horizon/escape edge
        pixels into the RGBA16F plane (shadow_coverage.wgsl:1-136)."""
        k = target.scene if target.scene is not None else 0
        out = target.c_struct()
        _capi.check(self._lib.gl_trace_refine_shadow(self._handle, k, C.byref(out),
                                                     _stream_handle(stream)))

    def shadow_edges(self, stream=None) -> int:
        """Edge pixels listed by the last refine_shadow (synchronizes)."""
        n = C.c_uint64()
        _capi.check(self._lib.gl_trace_plan_shadow_edges(self._handle, C.byref(n),
                                                         _stream_handle(stream)))
        return n.value

    def surface_footprint(self, scene: int = 0, pixels=None, stream=None) -> dict:
        """Source-chart Jacobians d(r, r_c dphi)/d(pixel x, y) of surface
        pixels by +-0.25-pixel central differences of real traces
        (surface_footprint_v1, footprint.rs:74-151). `pixels`: indices
        y * width + x (default: every pixel). Returns numpy arrays: resolved,
        parity (1 positive, 2 negative, 3 degenerate), J (n, 2, 2) with rows
        (radial, local arc) and columns (x, y), singular values (n, 2)."""
        n_all = self.scene_rays(scene)
        if pixels is None:
            d_pix, n = None, n_all
        else:
            d_pix = torch.as_tensor(np.asarray(pixels, dtype=np.int32), device="cuda")
            n = int(d_pix.numel())
        jac = torch.zeros(4 * max(n, 1), dtype=torch.float32, device="cuda")
        meta = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
        _capi.check(self._lib.gl_trace_surface_footprint(
            self._handle, scene, None if d_pix is None else d_pix.data_ptr(), n, jac.data_ptr(),
            meta.data_ptr(), _stream_handle(stream)))
        torch.cuda.synchronize()
        return _footprint_result(jac, meta, n)

    def set_persistent_config(self, block_size=256, blocks_per_cu=2, refill_interval=8,
                              refill_threshold=64, waves_per_simd=2) -> None:
        """Launch parameters of the persistent kernel (gl_persistent_config);
        the defaults are a new plan's: whole-tile mode, two waves per SIMD."""
        c = _capi.PersistentConfig(block_size, blocks_per_cu, refill_interval, refill_threshold,
                                   waves_per_simd)
        _capi.check(self._lib.gl_trace_plan_set_persistent_config(self._handle, C.byref(c)))

    def set_f64_overlap(self, enable: bool) -> None:
        """f64 overlap for frame sequences (gl_trace_plan_set_f64_overlap): a
        persistent launch's f64 passes run on the plan's second stream beside
        the next launch's trace; its outputs are complete after the next
        launch's trace kernel or after flush(). Consecutive launches must
        write different targets."""
        _capi.check(self._lib.gl_trace_plan_set_f64_overlap(self._handle, 1 if enable else 0))

    def flush(self, stream=None) -> None:
        """Makes `stream` wait for every f64 pass still in flight on the
        plan's second stream (gl_trace_plan_flush)."""
        _capi.check(self._lib.gl_trace_plan_flush(self._handle, _stream_handle(stream)))

    def persistent_config(self) -> dict:
        c = _capi.PersistentConfig()
        _capi.check(self._lib.gl_trace_plan_get_persistent_config(self._handle, C.byref(c)))
        return {k: getattr(c, k) for k, _ in c._fields_}

    def queue_stats(self, stream=None) -> dict:
        s = _capi.QueueStats()
        _capi.check(self._lib.gl_trace_plan_queue_stats(self._handle, C.byref(s),
                                                        _stream_handle(stream)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def encode_work(self, items: list, stream=None) -> None:
        """One persistent launch over a mixed work list. Each item is
        ("image", scene, TraceTarget) or ("batch", scene, d_states (n, 8) f64,
        d_dirs (n,) f32, TraceTarget), device tensors kept alive by the
        caller until the launch completes."""
        arr = (_capi.WorkItem * len(items))()
        for i, it in enumerate(items):
            w = arr[i]
            if it[0] == "image":
                w.scene, w.kind, w.count = it[1], _capi.WORK_IMAGE, 0
                w.outputs = it[2].c_struct()
            elif it[0] == "batch":
                _, scene, d_states, d_dirs, tgt = it
                w.scene, w.kind, w.count = scene, _capi.WORK_GEODESIC_BATCH, d_states.shape[0]
                w.states8, w.directions = d_states.data_ptr(), d_dirs.data_ptr()
                w.outputs = tgt.c_struct()
            else:
                raise ValueError(f"unknown work item kind {it[0]!r}")
        _capi.check(self._lib.gl_trace_encode_work(self._handle, arr, len(items),
                                                   _stream_handle(stream)))

    def enable_timing(self, enable: bool = True) -> None:
        """HIP events around the main kernel and the f64 finisher of each launch."""
        _capi.check(self._lib.gl_trace_plan_enable_timing(self._handle, int(enable)))

    def kernel_ms(self) -> tuple[float, float, int]:
        """(main kernel ms, f64 finisher ms, launches): mean durations over every
        launch since timing was enabled or this was last called."""
        a, b, n = C.c_double(), C.c_double(), C.c_uint64()
        _capi.check(self._lib.gl_trace_plan_kernel_ms(self._handle, C.byref(a), C.byref(b),
                                                      C.byref(n)))
        return a.value, b.value, int(n.value)

    def frame_graph(self, targets, scenes=None, stream=None) -> "FrameGraph":
        return FrameGraph(self, targets, scenes, stream)


class FrameGraph:
    """The persistent frame (queue reset + launch) captured once as a hipGraph."""

    def __init__(self, pipeline: TracePipeline, targets, scenes=None, stream=None):
        self._lib = pipeline._lib
        self.pipeline = pipeline
        if isinstance(targets, TraceTarget):
            targets = [targets]
        self.targets = targets
        outs = (_capi.RayOutputsC * len(targets))(*[t.c_struct() for t in targets])
        if scenes is None:
            ids, n = None, len(pipeline.scenes)
        else:
            self._ids = (C.c_uint32 * len(scenes))(*scenes)
            ids, n = C.cast(self._ids, C.c_void_p), len(scenes)
        self.stream = stream if stream is not None else torch.cuda.Stream()
        h = C.c_void_p()
        with torch.cuda.stream(self.stream):
            _capi.check(self._lib.gl_frame_graph_create(pipeline.handle, ids, n, outs,
                                                        _stream_handle(self.stream), C.byref(h)))
        self._handle = h

    def launch(self) -> None:
        _capi.check(self._lib.gl_frame_graph_launch(self._handle, _stream_handle(self.stream)))

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._lib.gl_frame_graph_destroy(h)
            self._handle = None


class GeodesicBatchError(ValueError):
    """Invalid batch configuration (GeodesicBatchError, batch.rs:7-17)."""


class GeodesicBatch:
    """Ordered batch of geodesic traces on the GPU (GeodesicBatch::trace_ordered).

    ``trace_ordered(states, directions)`` takes n initial states
    (t, x, y, z, p_t, p_x, p_y, p_z) in f64 and affine directions (+1 forward,
    -1 backward); the spacetime, policy and events are those of scene
    `scene` of the pipeline. Outcome i belongs to request i.

    ``threads`` mirrors GeodesicBatch::new(NonZeroUsize) (batch.rs:19-29): it
    must lie in [1, 256] as there, and is otherwise not used, since one GPU
    lane traces each request.
    """

    MAX_THREADS = 256

    def __init__(self, pipeline: TracePipeline, scene: int = 0, threads: int = 1):
        if not isinstance(threads, int) or threads < 1 or threads > self.MAX_THREADS:
            raise GeodesicBatchError(
                f"thread count must lie in [1, {self.MAX_THREADS}], got {threads!r}")
        self.pipeline = pipeline
        self.scene = scene
        self.threads = threads

    # requests per call of the allocation-free entry point
    NO_ALLOC_MAX_REQUESTS = 32768

    def trace_ordered(self, states: np.ndarray, directions: np.ndarray, stream=None,
                      persistent: bool = False, config: dict | None = None,
                      no_alloc: bool = False, f64_tolerance: tuple | None = None) -> dict:
        """Host arrays in, decoded host outputs back (synchronizes)."""
        states = np.ascontiguousarray(states, dtype=np.float64).reshape(-1, 8)
        n = states.shape[0]
        dirs = np.ascontiguousarray(directions, dtype=np.float32).reshape(-1)
        if dirs.shape[0] != n:
            raise ValueError("one direction per state is required")
        # ReferenceOutcome carries the turning radius (outcome.rs:425-436)
        target = TraceTarget(n, turning_radius=True)
        if n == 0:
            return target.to_host()
        d_states = torch.from_numpy(states).cuda()
        d_dirs = torch.from_numpy(dirs).cuda()
        self.encode(d_states, d_dirs, target, stream, persistent, config, no_alloc, f64_tolerance)
        torch.cuda.synchronize()
        return target.to_host()

    def encode(self, d_states: torch.Tensor, d_dirs: torch.Tensor, target: TraceTarget,
               stream=None, persistent: bool = False, config: dict | None = None,
               no_alloc: bool = False, f64_tolerance: tuple | None = None) -> None:
        """Device arrays in (float64 (n, 8) states, float32 (n,) directions),
        outputs into `target`; enqueued on `stream`, no synchronization.
        ``no_alloc`` selects gl_geodesic_batch_trace_noalloc (no device
        allocation; at most NO_ALLOC_MAX_REQUESTS requests per call);
        ``f64_tolerance = (rtol, atol)`` gl_geodesic_batch_trace_f64 (every
        request in f64 at that tolerance)."""
        n = d_states.shape[0]
        out = target.c_struct()
        lib = self.pipeline._lib
        if f64_tolerance is not None:
            if persistent or no_alloc:
                raise ValueError("the f64 batch has its own kernel")
            rtol, atol = f64_tolerance
            _capi.check(lib.gl_geodesic_batch_trace_f64(
                self.pipeline.handle, self.scene, C.c_void_p(d_states.data_ptr()),
                C.c_void_p(d_dirs.data_ptr()), n, float(rtol), float(atol), C.byref(out),
                _stream_handle(stream)))
        elif no_alloc:
            if persistent:
                raise ValueError("the allocation-free batch runs the plain batch kernel")
            _capi.check(lib.gl_geodesic_batch_trace_noalloc(
                self.pipeline.handle, self.scene, C.c_void_p(d_states.data_ptr()),
                C.c_void_p(d_dirs.data_ptr()), n, C.byref(out), _stream_handle(stream)))
        elif persistent:
            cfg = None
            if config is not None:
                cfg = C.byref(_capi.PersistentConfig(
                    config.get("block_size", 256), config.get("blocks_per_cu", 2),
                    config.get("refill_interval", 8), config.get("refill_threshold", 16),
                    config.get("waves_per_simd", 2)))
            _capi.check(lib.gl_geodesic_batch_trace_persistent(
                self.pipeline.handle, self.scene, C.c_void_p(d_states.data_ptr()),
                C.c_void_p(d_dirs.data_ptr()), n, cfg, C.byref(out), _stream_handle(stream)))
        else:
            _capi.check(lib.gl_geodesic_batch_trace(self.pipeline.handle, self.scene,
                                                    C.c_void_p(d_states.data_ptr()),
                                                    C.c_void_p(d_dirs.data_ptr()), n,
                                                    C.byref(out), _stream_handle(stream)))


def flop_constants() -> tuple[float, float]:
    """Counted flops of one RHS evaluation and of one accepted DP5 attempt."""
    lib = _capi.load()
    a, b = C.c_double(), C.c_double()
    _capi.check(lib.gl_flop_constants(C.byref(a), C.byref(b)))
    return a.value, b.value


class VjpTrace:
    """Discrete-adjoint vector-Jacobian product of the per-pixel frequency
    ratio g of one image scene (gl_vjp_forward / gl_vjp_backward).

    The constructor runs the forward trace (outputs in ``self.target``) and
    keeps the adjoint's saved state: the accepted step sizes, a checkpoint
    every ``checkpoint_interval`` accepted steps, and the terminal state.
    ``backward(grad_g)`` returns dL/d(a, Q, inclination) for
    L = sum_pixels grad_g * g over surface hits, and, if asked, the per-pixel
    gradient with respect to the sight-plane coordinates (s_x, s_y)."""

    def __init__(self, pipeline: TracePipeline, scene: int = 0, checkpoint_interval: int = 32,
                 stream=None):
        self.pipeline = pipeline
        self.scene = scene
        sc = pipeline.scenes[scene]
        self.target = pipeline.create_target(scene)
        self._handle = C.c_void_p()
        out = self.target.c_struct()
        _capi.check(pipeline._lib.gl_vjp_forward(pipeline.handle, scene,
                                                 math.radians(sc.inclination_deg),
                                                 checkpoint_interval, C.byref(out),
                                                 _stream_handle(stream), C.byref(self._handle)))

    @property
    def g(self) -> torch.Tensor:
        """Per-pixel g (0 where the ray did not end on the surface)."""
        st = self.target.status
        hit = (st & 15) == 7
        n = self.target.rays
        return torch.where(hit, self.target.observable[2 * n:], torch.zeros_like(self.target.travel_time))

    def backward(self, grad_g: torch.Tensor, direction: bool = False, stream=None):
        grad_g = grad_g.detach().to(device=self.target.status.device, dtype=torch.float32).contiguous()
        if grad_g.numel() != self.target.rays:
            raise ValueError("grad_g needs one value per pixel")
        gp = (C.c_double * 3)()
        gdir = torch.zeros(2 * self.target.rays, dtype=torch.float32,
                           device=grad_g.device) if direction else None
        _capi.check(self.pipeline._lib.gl_vjp_backward(
            self._handle, C.c_void_p(grad_g.data_ptr()), gp,
            C.c_void_p(gdir.data_ptr()) if gdir is not None else None, _stream_handle(stream)))
        params = np.array([gp[0], gp[1], gp[2]], dtype=np.float64)
        if gdir is None:
            return params
        return params, gdir.view(2, -1)

    def backward_observables(self, grad_r=None, grad_phi=None, grad_g=None,
                             direction: bool = False, stream=None):
        """dL/d(a, Q, inclination) (and per-pixel dL/d(s_x, s_y)) for
        L = sum_pixels (grad_r r + grad_phi phi_s + grad_g g) over surface hits."""
        dev = self.target.status.device
        seeds = []
        for v in (grad_r, grad_phi, grad_g):
            if v is None:
                seeds.append(None)
                continue
            t = torch.as_tensor(v).detach().to(device=dev, dtype=torch.float32).contiguous()
            if t.numel() != self.target.rays:
                raise ValueError("a seed needs one value per pixel")
            seeds.append(t)
        gp = (C.c_double * 3)()
        gdir = torch.zeros(2 * self.target.rays, dtype=torch.float32, device=dev) if direction else None
        ptr = [C.c_void_p(t.data_ptr()) if t is not None else None for t in seeds]
        _capi.check(self.pipeline._lib.gl_vjp_backward_observables(
            self._handle, ptr[0], ptr[1], ptr[2], gp,
            C.c_void_p(gdir.data_ptr()) if gdir is not None else None, _stream_handle(stream)))
        params = np.array([gp[0], gp[1], gp[2]], dtype=np.float64)
        if gdir is None:
            return params
        return params, gdir.view(2, -1)

    def surface_footprint(self, stream=None) -> dict:
        """Source-chart footprint of every pixel from two adjoint passes (seed
        r, then phi_s; gl_vjp_surface_footprint): the analytic counterpart of
        TracePipeline.surface_footprint's +-0.25-pixel central differences
        (surface_footprint_v1, footprint.rs:74-151), same result layout."""
        n = self.target.rays
        dev = self.target.status.device
        jac = torch.zeros(4 * n, dtype=torch.float32, device=dev)
        meta = torch.zeros(n, dtype=torch.int32, device=dev)
        out = self.target.c_struct()
        _capi.check(self.pipeline._lib.gl_vjp_surface_footprint(
            self._handle, C.byref(out), C.c_void_p(jac.data_ptr()), C.c_void_p(meta.data_ptr()),
            _stream_handle(stream)))
        torch.cuda.synchronize()
        return _footprint_result(jac, meta, n)

    def lensing_jacobian(self) -> np.ndarray:
        """Per-pixel d(r, r phi_s)/d(pixel x, y), shape (n, 2, 2), zero where
        the footprint does not resolve."""
        return self.surface_footprint()["J"]

    def close(self) -> None:
        if self._handle:
            self.pipeline._lib.gl_vjp_destroy(self._handle)
            self._handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001  (interpreter shutdown)
            pass


class FrequencyRatioImage(torch.autograd.Function):
    """torch.autograd.Function over the adjoint trace: forward(params, scene)
    with params = (a, Q, inclination [rad]) returns the per-pixel g image of
    `scene` traced at those parameters; backward returns dL/dparams from the
    discrete adjoint (the scene object is not differentiated)."""

    @staticmethod
    def forward(ctx, params: torch.Tensor, scene: Scene, checkpoint_interval: int = 32):
        a, q, incl = (float(v) for v in params.detach().cpu().tolist())
        sc = Scene(**{**scene.__dict__, "spin": a, "charge": q,
                      "inclination_deg": math.degrees(incl)})
        pipe = TracePipeline([sc])
        vjp = VjpTrace(pipe, 0, checkpoint_interval)
        ctx.vjp = vjp
        ctx.pipe = pipe
        return vjp.g.clone()

    @staticmethod
    def backward(ctx, grad_g):
        gp = ctx.vjp.backward(grad_g)
        ctx.vjp.close()
        return torch.tensor(gp, dtype=torch.float64), None, None
