"""The scene checks of the C ABI (gl_scene_validate) accept exactly the
observers the reference accepts: the stationary-observer tetrad of
observer.rs:106-183 with its Gram residual measured against each product's
term norm (observer.rs:354-376, metric_dot_and_term_norm spacetime.rs:463-483)
and a 1e-12 budget, restated by the oracle (oracle/observer.cpp).

Every one of the first 300 draws of the uneven-scene stream (seed 20261206,
the 256-scene workload) builds in the oracle iff the C ABI validates it,
including two draws that an absolute Gram residual refused:
(a, Q, inclination, r_obs) = (0.082, 0.782, 74.8 deg, 196.3 M) and
(0.062, 0.371, 8.5 deg, 67.2 M)."""
import ctypes as C
import math

import pytest

from gravlume_amd import _capi, scene as S


def _oracle_accepts(oracle, s: S.Scene) -> bool:
    try:
        oracle.observation(1.0, s.spin, s.charge, float(s.chart), [0.0, *s.observer_xyz()],
                           [0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 1.0], 1.0, s.width, s.height,
                           math.radians(s.fov_deg))
        return True
    except ValueError:
        return False


def test_uneven_draws_validate_iff_the_oracle_builds_them(oracle):
    lib = _capi.load()
    draws = S.uneven_scene_draws(20261206)
    agree, refused = 0, []
    for i in range(300):
        s = next(draws)
        d = s.to_desc()
        ours = lib.gl_scene_validate(C.byref(d)) == 0
        theirs = _oracle_accepts(oracle, s)
        assert ours == theirs, (i, s.spin, s.charge, s.inclination_deg, s.observer_radius,
                                lib.gl_last_error())
        agree += 1
        if not ours:
            refused.append(i)
    assert agree == 300
    # the workload is the plain draw sequence
    assert refused == [] or min(refused) >= 256


@pytest.mark.parametrize("a,q,incl,r_obs", [(0.082, 0.782, 74.8, 196.3), (0.062, 0.371, 8.5, 67.2)])
def test_strong_metric_observers_validate(oracle, a, q, incl, r_obs):
    s = S.Scene(spin=a, charge=q, inclination_deg=incl, observer_radius=r_obs, width=256,
                height=256, fov_deg=20.0, disk=(None, 20.0), escape_map=False)
    d = s.to_desc()
    assert _capi.load().gl_scene_validate(C.byref(d)) == 0, _capi.load().gl_last_error()
    assert _oracle_accepts(oracle, s)
