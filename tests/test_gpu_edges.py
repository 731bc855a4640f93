"""Edge cases of the C ABI on the GPU (needs an MI355X): empty requests,
the smallest viewports, output planes too small, pixel lists that leave the
image, and the errors the host checks raise (GpuTraceInputError's role,
crates/gravlume-render/src/trace/input.rs:342-381)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gl():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gravlume_amd
    return gravlume_amd


def _default_view(S, **kw):
    # the reference's default view (gpu_trace_tests.rs:849-855: a = 0.8, the
    # outgoing chart, r = 30 M at 60 deg) with a [6, 20] M surface
    base = dict(spin=0.8, charge=0.0, inclination_deg=60.0, observer_radius=30.0, width=64,
                height=36, fov_deg=45.0, chart=S.OUTGOING, disk=(6.0, 20.0), escape_radius=200.0,
                escape_map=False, emission="bolometric")
    base.update(kw)
    return S.Scene(**base)


def test_empty_requests_are_no_ops(gl):
    """An empty geodesic batch (batch.rs:31-43) and an empty footprint list
    return OK and launch nothing, in the plain and the persistent batch."""
    from gravlume_amd import scene as S
    import torch
    pipe = gl.TracePipeline([_default_view(S)])
    b = gl.GeodesicBatch(pipe, 0)
    for persistent in (False, True):
        out = b.trace_ordered(np.zeros((0, 8)), np.zeros(0), persistent=persistent)
        assert out["termination"].shape == (0,)
    lib = pipe._lib
    t = gl.TraceTarget(1)
    st = torch.zeros(8, dtype=torch.float64, device="cuda")
    dr = torch.ones(1, dtype=torch.float32, device="cuda")
    gl._capi.check(lib.gl_geodesic_batch_trace(pipe.handle, 0, C.c_void_p(st.data_ptr()),
                                               C.c_void_p(dr.data_ptr()), 0, C.byref(t.c_struct()),
                                               None))
    jac = torch.zeros(4, dtype=torch.float32, device="cuda")
    meta = torch.zeros(1, dtype=torch.int32, device="cuda")
    pix = torch.zeros(1, dtype=torch.int32, device="cuda")
    gl._capi.check(lib.gl_trace_surface_footprint(pipe.handle, 0, pix.data_ptr(), 0,
                                                  jac.data_ptr(), meta.data_ptr(), None))
    torch.cuda.synchronize()
    assert int(meta[0]) == 0


@pytest.mark.parametrize("wh", [(1, 1), (1, 9), (9, 1), (7, 5)])
def test_smallest_viewports_trace_alike_in_every_dispatch(gl, oracle, wh):
    """Viewports of a single pixel, a single row or column, and a partial
    tile: the plain, whole-tile and lane-refill dispatches agree bit for bit,
    and every definite class matches the oracle's except where the oracle's
    own class changes under a 16x tighter tolerance."""
    import math
    from gravlume_amd import scene as S
    import torch
    w, h = wh
    sc = _default_view(S, width=w, height=h, fov_deg=30.0)
    pipe = gl.TracePipeline([sc])
    a, b, c = pipe.create_target(0), pipe.create_target(0), pipe.create_target(0)
    pipe.encode_batch(a)
    pipe.set_persistent_config(block_size=256, blocks_per_cu=2, refill_interval=8,
                               refill_threshold=16, waves_per_simd=2)
    pipe.encode_persistent([b])
    pipe.set_persistent_config(block_size=256, blocks_per_cu=2, refill_interval=8,
                               refill_threshold=64, waves_per_simd=2)
    pipe.encode_persistent([c])
    torch.cuda.synchronize()
    ha, hb, hc = a.to_host(), b.to_host(), c.to_host()
    for k in ha:
        assert np.array_equal(np.asarray(ha[k]), np.asarray(hc[k])), k
        assert np.array_equal(np.asarray(ha[k]), np.asarray(hb[k])), k
    obs = oracle.observation(1.0, sc.spin, sc.charge, float(sc.chart), [0.0, *sc.observer_xyz()],
                             [0, 0, 0, 0], [0, 0, 1], 1.0, w, h, math.radians(sc.fov_deg))
    ev = oracle.events(sc.escape_radius, sc.disk_radii(), horizon_offset=sc.horizon_offset)
    idx = np.arange(w * h)
    sxy = oracle.sight_planes(obs, idx % w, idx // w)
    ref, _ = oracle.observation_trace_ordered(obs, oracle.config_policy(1e-10, 1e-12), ev, sxy,
                                              oracle.Emitter(), threads=4)
    fine, _ = oracle.observation_trace_ordered(obs, oracle.config_policy(1e-10 / 16, 1e-12 / 16),
                                               ev, sxy, oracle.Emitter(), threads=4)
    A, F = oracle.outcomes_to_arrays(ref), oracle.outcomes_to_arrays(fine)
    settled = A["termination"] == F["termination"]  # not near-critical for the oracle itself
    assert settled.sum() >= len(idx) - 1
    definite = settled & (ha["termination"] != 6)
    assert definite.sum() >= len(idx) - 1
    assert np.array_equal(ha["termination"][definite], A["termination"][definite])


def test_output_planes_too_small_are_rejected(gl):
    from gravlume_amd import scene as S
    pipe = gl.TracePipeline([_default_view(S)])
    small = gl.TraceTarget(63 * 36)
    with pytest.raises(gl._capi.GravlumeError) as e:
        pipe.encode_batch(small)
    assert e.value.status == gl._capi.GL_ERR_INVALID_ARGUMENT
    with pytest.raises(gl._capi.GravlumeError):
        pipe.encode_persistent([small])


def test_footprint_pixel_outside_the_image_is_not_traced(gl):
    """A pixel list entry beyond width x height gives zeros and meta 0; the
    entries around it are traced as usual."""
    from gravlume_amd import scene as S
    sc = _default_view(S, width=128, height=72)
    pipe = gl.TracePipeline([sc])
    inside = [36 * 128 + 64, 35 * 128 + 90]
    fp = pipe.surface_footprint(0, [inside[0], 128 * 72, inside[1], 2 ** 31 - 5])
    assert not fp["resolved"][1] and not fp["resolved"][3]
    assert fp["termination"][1] == 0 and fp["termination"][3] == 0
    assert not fp["J"][[1, 3]].any()
    ref = pipe.surface_footprint(0, inside)
    assert np.array_equal(fp["J"][[0, 2]], ref["J"])
    assert np.array_equal(fp["termination"][[0, 2]], ref["termination"])


def test_whole_tile_launches_in_adapted_tile_order_stay_bit_identical(gl):
    """Each whole-tile launch hands out its tiles by decreasing cost of the
    previous launch (the order kernel after the launch): four consecutive
    launches over a two-scene table, and over a hipGraph of the same frame,
    all equal the plain dispatch bit for bit."""
    from gravlume_amd import scene as S
    import torch
    scenes = [S.single_frame_1080p(width=480, height=270),
              _default_view(S, width=200, height=120)]
    pipe = gl.TracePipeline(scenes)
    refs = []
    for i in range(2):
        t = pipe.create_target(i)
        pipe.encode_batch(t)
        refs.append(t)
    torch.cuda.synchronize()
    refs = [t.to_host() for t in refs]
    pipe.set_persistent_config(block_size=256, blocks_per_cu=2, refill_interval=8,
                               refill_threshold=64, waves_per_simd=2)
    ts = [pipe.create_target(0), pipe.create_target(1)]
    for launch in range(4):
        for t in ts:
            t.status.zero_()
        pipe.encode_persistent(ts)
        torch.cuda.synchronize()
        for i in range(2):
            h = ts[i].to_host()
            for k in refs[i]:
                assert np.array_equal(np.asarray(h[k]), np.asarray(refs[i][k])), (launch, i, k)
    graph = pipe.frame_graph(ts)
    for _ in range(3):
        for t in ts:
            t.status.zero_()
        torch.cuda.synchronize()
        graph.launch()
        graph.stream.synchronize()
        for i in range(2):
            h = ts[i].to_host()
            for k in refs[i]:
                assert np.array_equal(np.asarray(h[k]), np.asarray(refs[i][k])), ("graph", i, k)
