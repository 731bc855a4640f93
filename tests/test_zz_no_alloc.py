"""gl_geodesic_batch_trace_noalloc (needs an MI355X): the allocation-free
geodesic batch gives gl_geodesic_batch_trace's outputs bit for bit, on two
calls in a row of which the second is smaller (its static scratch keeps the
first call's records), on a scene that uses both f64 passes (near-horizon
deferral and the near-critical refine pass); it refuses more than 32768
requests with GL_ERR_CAPACITY before launching anything.
"""
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


def _states(n, seed):
    from gravlume_amd import critical as Cr, scene as S
    sc = S.near_critical_scene()
    states, *_ = Cr.near_critical_states(sc.spin, sc.charge, sc.observer_radius,
                                         math.radians(sc.inclination_deg), n, seed=seed,
                                         chart=sc.chart)
    return states


def test_no_alloc_batch_matches_the_allocating_entry_point(gl):
    from gravlume_amd import scene as S
    pipe = gl.TracePipeline([S.near_critical_scene(refine=True)])
    batch = gl.GeodesicBatch(pipe, 0)
    for n, seed in ((4096, 20261210), (1000, 20261211)):
        states = _states(n, seed)
        dirs = -np.ones(n)
        print("no-alloc batch: requests", n, "seed", seed, flush=True)
        want = batch.trace_ordered(states, dirs)
        have = batch.trace_ordered(states, dirs, no_alloc=True)
        assert np.all(have["termination"] > 0)
        assert have["refined"].any() and (have["termination"] == 1).any()
        for k in want:
            assert np.array_equal(np.asarray(have[k]), np.asarray(want[k]), equal_nan=True), k


def test_no_alloc_batch_refuses_more_than_its_capacity(gl):
    import torch
    from gravlume_amd import scene as S
    pipe = gl.TracePipeline([S.near_critical_scene()])
    n = gl.GeodesicBatch.NO_ALLOC_MAX_REQUESTS + 1
    d_states = torch.zeros((n, 8), dtype=torch.float64, device="cuda")
    d_dirs = -torch.ones(n, dtype=torch.float32, device="cuda")
    tgt = gl.TraceTarget(n)
    with pytest.raises(gl._capi.GravlumeError) as e:
        gl.GeodesicBatch(pipe, 0).encode(d_states, d_dirs, tgt, no_alloc=True)
    assert e.value.status == gl._capi.GL_ERR_CAPACITY
