"""Near-critical initial data (gravlume_amd/critical.py) against the oracle:
the f64 states carry exactly the requested constants of motion (E = 1,
L_z = xi, Carter = eta) and are null; circular photon orbits reduce to the
Kerr closed form (Bardeen, Press & Teukolsky 1972) at Q = 0; and rays just
outside the critical curve escape while rays just inside fall in."""
import math

import numpy as np


def test_states_carry_the_requested_constants(oracle):
    from gravlume_amd import critical as Cr
    a, q = 0.998, 0.05
    states, delta, xi, eta = Cr.near_critical_states(a, q, 200.0, math.radians(85.0), 64, seed=1)
    st = oracle.spacetime(1.0, a, q, -1.0)
    for i in range(len(states)):
        e, lz, carter, null = oracle.invariants(st, states[i])
        assert abs(e - 1.0) <= 1e-13
        assert abs(lz - xi[i]) <= 1e-12 * max(1.0, abs(xi[i]))
        assert abs(carter - eta[i]) <= 1e-11 * max(1.0, abs(eta[i]))
        assert null <= 1e-13


def test_circular_photon_orbits_match_the_kerr_closed_form():
    from gravlume_amd import critical as Cr
    for a in (0.3, 0.8, 0.998):
        r1, r2 = Cr.circular_photon_orbits(a, 0.0)
        want1 = 2.0 * (1.0 + math.cos(2.0 / 3.0 * math.acos(-a)))
        want2 = 2.0 * (1.0 + math.cos(2.0 / 3.0 * math.acos(a)))
        assert abs(r1 - want1) <= 1e-9 and abs(r2 - want2) <= 1e-9
    # Reissner-Nordstrom photon sphere
    assert abs(Cr.photon_sphere(0.0) - 3.0) <= 1e-15


def test_offsets_split_escape_and_capture(oracle):
    from gravlume_amd import critical as Cr
    a, q = 0.998, 0.05
    states, delta, _, _ = Cr.near_critical_states(a, q, 200.0, math.radians(85.0), 48, seed=2,
                                                  delta_range=(1e-4, 1e-3))
    st = oracle.spacetime(1.0, a, q, -1.0)
    ev = oracle.events(1000.0, None, horizon_offset=1e-6)
    outs = oracle.trace_ordered(st, oracle.config_policy(1e-12, 1e-14), ev, states,
                                -np.ones(len(states)), threads=8)
    A = oracle.outcomes_to_arrays(outs)
    assert np.all((A["termination"] == 2) == (delta > 0))
    assert np.all((A["termination"] == 1) == (delta < 0))
