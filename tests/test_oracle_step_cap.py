"""The oracle's optional relative step cap (oracle/trace.cpp step_cap), used
only by the CPU baseline that takes the GPU trace's step limits.

* With the cap off (0, the reference policy's value) the trace is the
  reference policy's, bit for bit: the same outcome fields as the preset.
* With |h| <= 0.25 r and a 256 M maximum step (the GPU trace's limits), the
  largest accepted step of a ray that escapes to r = 1000 stays within a
  quarter of the escape radius, the ray takes a small fraction of the
  0.5 M-capped policy's steps, and class, escape direction and travel time
  agree with the reference policy's trace within the tolerance's reach.
"""
import math

import numpy as np
import pytest


def _frame_rays(oracle, n=64):
    obs = oracle.observation(1.0, 0.94, 0.0, -1.0, [0.0, *_observer_xyz()], [0, 0, 0, 0],
                             [0, 0, 1], 1.0, 1920, 1080, math.radians(15.0))
    idx = np.random.default_rng(3).choice(1920 * 1080, n, replace=False)
    return obs, oracle.sight_planes(obs, idx % 1920, idx // 1920)


def _observer_xyz():
    from gravlume_amd import scene as S
    return S.single_frame_1080p().observer_xyz()


def _trace(oracle, obs, sxy, pol):
    ev = oracle.events(1000.0, None)
    outs, _ = oracle.observation_trace_ordered(obs, pol, ev, sxy, None, 4)
    return oracle.outcomes_to_arrays(outs)


def test_cap_off_is_the_reference_policy(oracle):
    obs, sxy = _frame_rays(oracle, 16)
    a = _trace(oracle, obs, sxy, oracle.config_policy(1e-6, 1e-8))
    p = oracle.config_policy(1e-6, 1e-8)
    p.maximum_step_per_radius = 0.0
    b = _trace(oracle, obs, sxy, p)
    for k in ("termination", "accepted", "rejected", "travel_time", "affine", "state"):
        assert np.array_equal(a[k], b[k]), k


def test_relative_cap_bounds_steps_and_keeps_the_observables(oracle):
    obs, sxy = _frame_rays(oracle, 64)
    ref = _trace(oracle, obs, sxy, oracle.config_policy(1e-10, 1e-12))
    gpu = _trace(oracle, obs, sxy, oracle.gpu_step_policy(1e-10, 1e-12, 1.0e-2, 256.0, 0.25))
    esc = (ref["termination"] == 2) & (gpu["termination"] == 2)
    assert esc.sum() >= 16
    assert np.array_equal(ref["termination"], gpu["termination"])
    # the largest step of a ray that reaches r = 1000 is bounded by 0.25 * 1000 (and far above
    # the reference policy's 0.5 M)
    assert np.all(gpu["max_step"][esc] <= 0.25 * 1000.0 * 1.001)
    assert np.all(gpu["max_step"][esc] > 10.0)
    assert np.all(gpu["accepted"][esc] * 3 < ref["accepted"][esc])
    cosang = np.sum(gpu["escape_direction"][esc] * ref["escape_direction"][esc], axis=1)
    assert np.all(np.arccos(np.clip(cosang, -1.0, 1.0)) < 1e-6)
    assert np.allclose(gpu["travel_time"][esc], ref["travel_time"][esc], rtol=1e-7)
