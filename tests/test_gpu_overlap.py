"""f64 overlap for frame sequences (gl_trace_plan_set_f64_overlap; needs an
MI355X). The f64 passes of a persistent launch (near-horizon finisher,
near-critical refine) run on the plan's second stream beside the next
launch's trace. Every output must be bit-identical to the serial launch, and
complete on the caller's stream after the next launch, after flush, or after
any other entry point of the plan; every target is zeroed before it is
written again, so a launch that left a plane unwritten would show."""
import math

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


def _same(a, b, tag):
    for k in b:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (tag, k)


def _zero(t):
    for name in t.PLANES:
        plane = getattr(t, name)
        if plane is not None:
            plane.zero_()


def test_alternating_frames_match_serial(gl):
    import torch
    from gravlume_amd import scene as S
    sc = S.single_frame_1080p(width=512, height=288)
    pipe = gl.TracePipeline([sc])
    ref_t = pipe.create_target(0)
    pipe.encode_persistent([ref_t])
    torch.cuda.synchronize()
    ref = ref_t.to_host()
    assert np.any(ref["termination"] == 7)
    pipe.set_f64_overlap(True)
    A, B = pipe.create_target(0), pipe.create_target(0)
    targets = (A, B)
    for i in range(6):
        t = targets[i % 2]
        _zero(t)
        pipe.encode_persistent([t])
        if i >= 1:
            # the previous frame is complete on this stream once this
            # launch's trace is enqueued: read it without a device sync
            _same(targets[(i - 1) % 2].to_host(), ref, ("previous", i))
    pipe.flush()
    _same(targets[5 % 2].to_host(), ref, "flushed")
    # another entry point joins the outstanding passes: a plain launch
    _zero(A)
    pipe.encode_persistent([A])
    C = pipe.create_target(0)
    pipe.encode_batch(C)
    _same(A.to_host(), ref, "after plain launch")
    _same(C.to_host(), ref, "plain")
    # disabling the mode drains it
    _zero(B)
    pipe.encode_persistent([B])
    pipe.set_f64_overlap(False)
    _same(B.to_host(), ref, "after disable")
    torch.cuda.synchronize()


def test_refined_batches_and_graph_capture_under_overlap(gl):
    """Consecutive persistent near-critical batches with the f64 refine pass
    (the long f64 pass) into two targets, then a frame graph captured and
    replayed while the mode is on (capture drains the second stream and the
    captured launch runs its passes in line)."""
    import torch
    from gravlume_amd import critical as Cr, scene as S
    sc = S.near_critical_scene(refine=True)
    n = 20_000
    states, *_ = Cr.near_critical_states(sc.spin, sc.charge, sc.observer_radius,
                                         math.radians(sc.inclination_deg), n, seed=31,
                                         chart=sc.chart)
    pipe = gl.TracePipeline([sc])
    batch = gl.GeodesicBatch(pipe, 0)
    ref = batch.trace_ordered(states, -np.ones(n), persistent=True)
    assert ref["refined"].mean() > 0.01
    pipe.set_f64_overlap(True)
    d_states = torch.from_numpy(states).cuda()
    d_dirs = -torch.ones(n, dtype=torch.float32, device="cuda")
    outs = (gl.TraceTarget(n, turning_radius=True), gl.TraceTarget(n, turning_radius=True))
    for i in range(4):
        t = outs[i % 2]
        _zero(t)
        batch.encode(d_states, d_dirs, t, persistent=True)
    pipe.flush()
    for t in outs:
        _same(t.to_host(), ref, "batch")

    fsc = S.single_frame_1080p(width=256, height=144)
    fpipe = gl.TracePipeline([fsc])
    eager = fpipe.create_target(0)
    fpipe.encode_persistent([eager])
    torch.cuda.synchronize()
    he = eager.to_host()
    fpipe.set_f64_overlap(True)
    other = fpipe.create_target(0)
    fpipe.encode_persistent([other])  # leaves passes in flight on the second stream
    tgt = fpipe.create_target(0)
    graph = fpipe.frame_graph([tgt])
    for rep in range(3):
        _zero(tgt)
        torch.cuda.synchronize()
        graph.launch()
        graph.stream.synchronize()
        _same(tgt.to_host(), he, ("replay", rep))
        fpipe.encode_persistent([other if rep % 2 else eager])
        torch.cuda.synchronize()
    _same(other.to_host(), he, "eager between replays")
    _same(eager.to_host(), he, "eager between replays")
