"""Closed forms, invariants and special cases that pin the oracle beyond the
fixtures (crates/gravlume-domain/tests/spacetime_tests.rs,
crates/gravlume-reference/tests/geodesic_tests.rs, surface.rs:336-363,
spacetime.rs:780-904). Each check is chosen so that a dropped term, a wrong
sign or index in the restated formulas fails it."""
import math

import numpy as np
import pytest


def test_extremality_and_horizon_without_clamping(oracle):
    O = oracle
    f = np.finfo(float)
    assert O.spacetime(1, 0.8, 0).extremality == 0 and abs(O.spacetime(1, 0.8, 0).outer_horizon - 1.6) <= 2e-15
    assert O.spacetime(1, 1, 0).extremality == 1 and O.spacetime(1, 1, 0).outer_horizon == 1.0
    assert O.spacetime(1, 1, 0.5).extremality == 2 and not O.spacetime(1, 1, 0.5).has_horizon
    for m in (f.max / 2, f.tiny):
        st = O.spacetime(m, m, 0)
        assert st.extremality == 1 and st.outer_horizon == m
    below_one = np.nextafter(1.0, 0.0)
    assert O.spacetime(1, below_one, 2.0 ** -26).extremality == 2
    st = O.spacetime(1, below_one, 0)
    assert st.extremality == 0 and st.outer_horizon == 2.0 ** -26 + 1.0


@pytest.mark.parametrize("a,q", [(0.5, 0.3), (0.94, 0.0), (0.9, 0.2), (0.998, 0.05)])
def test_outer_horizon_closed_form(oracle, a, q):
    assert abs(oracle.spacetime(1, a, q).outer_horizon - (1 + math.sqrt(1 - a * a - q * q))) <= 4e-16


def test_schwarzschild_rn_minkowski_limits(oracle):
    O = oracle
    st = O.spacetime(2.0, 0.0, 0.0)
    g = O.geometry(st, [3.0, 4.0, 12.0])
    assert abs(g["radius"] - 13.0) <= np.finfo(float).eps * 13
    cov, _ = O.metric(st, [3.0, 4.0, 12.0])
    assert abs(cov[0, 0] - (-9.0 / 13.0)) <= 2e-15
    cov, _ = O.metric(O.spacetime(1.0, 0.0, 0.6), [10.0, 0.0, 0.0])
    assert abs(cov[0, 0] - (-1 + 0.2 - 0.36 / 100)) <= 4 * np.finfo(float).eps
    cov, inv = O.metric(O.spacetime(1e-12, 0.0, 0.0), [10.0, 0.0, 0.0])
    assert abs(cov[0, 0] + 1.0) <= 2.1e-13
    assert np.max(np.abs(cov @ inv - np.eye(4))) < 4 * np.finfo(float).eps


def test_far_field_and_axis_without_overflow(oracle):
    st = oracle.spacetime(1.0, 0.0, 0.0)
    for e in (80, 200, 300):
        x = 10.0 ** e
        assert oracle.geometry(st, [x, 0.0, 0.0])["radius"] == x
    st = oracle.spacetime(1.0, 1.0, 0.0)
    for e in (-300, -200, -160):
        z = 10.0 ** e
        assert oracle.geometry(st, [0.0, 0.0, z])["radius"] == z


def test_ring_and_branch_disk_are_distinct_failures(oracle):
    st = oracle.spacetime(1.0, 0.8, 0.0)
    with pytest.raises(ValueError, match="error 2"):
        oracle.geometry(st, [0.8, 0.0, 0.0])
    with pytest.raises(ValueError, match="error 3"):
        oracle.geometry(st, [0.0, 0.0, 0.0])
    with pytest.raises(ValueError, match="error 3"):
        oracle.geometry(st, [np.nextafter(0.8, 0.0), 0.0, 0.0])


@pytest.mark.parametrize("chart", [1.0, -1.0])
def test_both_charts_pull_back_to_boyer_lindquist(oracle, chart):
    # spacetime.rs:780-878: g(B_i, B_j) equals the BL metric for the chart basis.
    rng = np.random.default_rng(7)
    for _ in range(32):
        M = rng.uniform(0.5, 4.0)
        a = M * rng.uniform(-0.9, 0.9)
        q = M * rng.uniform(-0.2, 0.2)
        r = M * rng.uniform(3.0, 20.0)
        th = rng.uniform(0.1, math.pi - 0.1)
        ph = rng.uniform(-math.pi, math.pi)
        st = oracle.spacetime(M, a, q, chart)
        x, y, z = oracle.oblate_to_cartesian(a, chart, r, th, ph)
        cov, _ = oracle.metric(st, [x, y, z])
        s2 = math.sin(th) ** 2
        sigma = r * r + a * a * math.cos(th) ** 2
        delta = r * r - 2 * M * r + a * a + q * q
        rn = 2 * M * r - q * q
        rf = a * a + r * r
        exp = np.array([[-1 + rn / sigma, 0, 0, -rn * a * s2 / sigma],
                        [0, sigma / delta, 0, 0], [0, 0, sigma, 0],
                        [-rn * a * s2 / sigma, 0, 0, s2 * (rf * rf - a * a * delta * s2) / sigma]])
        cs = chart * a
        sp, cp = math.sin(ph), math.cos(ph)
        B = np.array([[1, 0, 0, 0],
                      [0, math.sin(th) * cp, math.sin(th) * sp, math.cos(th)],
                      [0, math.cos(th) * (r * cp - cs * sp), math.cos(th) * (r * sp + cs * cp), -r * math.sin(th)],
                      [0, -y, x, 0]], dtype=float)
        BL = B.copy()
        BL[1] = B[1] + chart * ((a / delta) * B[3] + (rn / delta) * B[0])
        got = BL @ cov @ BL.T
        scale = np.maximum(np.abs(BL) @ np.abs(cov) @ np.abs(BL).T, 1.0)
        assert np.all(np.abs(got - exp) <= 256 * np.finfo(float).eps * np.maximum(scale, np.abs(exp)))


def test_rhs_is_hamiltons_equations(oracle):
    # x' = dH/dp and p' = -dH/dx with H = g^{mn} p_m p_n / 2, by central differences.
    rng = np.random.default_rng(3)
    for chart in (1.0, -1.0):
        st = oracle.spacetime(1.0, 0.9, 0.3, chart)
        for _ in range(10):
            x = rng.uniform(-12, 12, 3)
            p = np.concatenate([[-1.0], rng.uniform(-1, 1, 3)])

            def H(xx, pp):
                _, inv = oracle.metric(st, xx)
                return 0.5 * pp @ inv @ pp

            d = oracle.rhs(st, np.concatenate([[0.0], x, p]))
            h = 1e-6
            for i in range(4):
                e = np.zeros(4); e[i] = h
                assert abs(d[i] - (H(x, p + e) - H(x, p - e)) / (2 * h)) <= 1e-7
            for i in range(3):
                e = np.zeros(3); e[i] = h
                assert abs(d[5 + i] + (H(x + e, p) - H(x - e, p)) / (2 * h)) <= 1e-7
            assert d[4] == 0.0


def test_axis_null_derivative_handedness(oracle):
    # spacetime.rs:880-904
    for chart in (1.0, -1.0):
        st = oracle.spacetime(1.0, 0.8, 0.0, chart)
        g = oracle.geometry(st, [0.0, 0.0, 4.0])
        e = 0.8 / (16 + 0.64)
        assert abs(g["null_vector_gradient"][0][2] + e) <= 8 * np.finfo(float).eps
        assert abs(g["null_vector_gradient"][1][1] - e) <= 8 * np.finfo(float).eps


def test_schwarzschild_circular_emitter_redshift(oracle):
    # surface.rs:336-363
    st = oracle.spacetime(1.0, 0.0, 0.0)
    em = oracle.Emitter()
    rc, out = oracle.observable_at(st, em, 6.0, 20.0, [0, 6, 0, 0, -1, 0, 0, 0], 1.0)
    assert rc == 0 and abs(out.frequency_ratio - math.sqrt(0.5)) <= 4 * np.finfo(float).eps
    rc, _ = oracle.observable_at(st, em, 3.0, 3.0, [0, 3, 0, 0, -1, 0, 0, 0], 1.0)
    assert rc == 4  # photon orbit is not a timelike emitter


def test_weak_field_deflection(oracle):
    # geodesic_tests.rs:73-110
    R, b = 1000.0, 50.0
    pt = b / R
    f = 2.0 / R
    quad = f * f - (1 - f) * (pt * pt - 1 - f)
    pr = (f - math.sqrt(quad)) / (1 - f)
    out = oracle.trace(oracle.spacetime(1, 0, 0), oracle.policy("regular"), oracle.events(R),
                       [0, R, 0, 0, -1, pr, pt, 0], 1.0)
    assert out.termination == oracle.TERM_ESCAPE
    flat = math.pi - 2 * math.asin(b / R)
    assert abs((out.azimuth_advance - flat) - 4.0 / b) <= 0.08 * 4.0 / b


def test_equatorial_surface_is_a_terminal_event(oracle):
    # geodesic_tests.rs:112-131
    out = oracle.trace(oracle.spacetime(1, 0, 0), oracle.policy("regular"),
                       oracle.events(None, surface=(5.0, 60.0), has_escape=False),
                       [0, 50, 0, 1, -1, -0.9, 0, -0.2], 1.0)
    assert out.termination == oracle.TERM_SURFACE and abs(out.state[3]) <= 2e-11


def test_single_step_exhaustion(oracle):
    # geodesic.rs:753-770
    pol = oracle.policy("regular", max_accepted=1)
    out = oracle.trace(oracle.spacetime(1, 0, 0), pol, oracle.events(None, has_escape=False),
                       [0, 50, 0, 0, -1, -0.99, 0.1, 0], 1.0)
    assert out.termination == oracle.TERM_STEP_EXHAUSTION and out.accepted == 1


def test_superextremal_ring_approach_hits_the_guard(oracle):
    # geodesic.rs:818-837
    out = oracle.trace(oracle.spacetime(1, 2.0, 0), oracle.policy("regular"),
                       oracle.events(None, has_escape=False), [0, 2.1, 0, 0, -1, -1, 0, 0], 1.0)
    assert out.termination == oracle.TERM_SINGULARITY


def test_invariants_conserved_on_a_kerr_newman_orbit(oracle):
    # E is exactly conserved by construction (p_t' = 0); L_z and the trig-form
    # Carter constant drift only at the integrator tolerance on a null ray.
    xyz = oracle.oblate_to_cartesian(0.9, -1.0, 30.0, 1.2, 0.3)
    obs = oracle.observation(1.0, 0.9, 0.2, -1.0, [0, *xyz], [0, 0, 0, 0], [0, 0, 1], 1.0, 64, 64,
                             0.5)
    for px, py in ((10, 20), (40, 33), (63, 0)):
        y, _, _, null = oracle.initial_ray(obs, *oracle.sight_plane(obs, px, py))
        assert null <= 2e-12
        out = oracle.trace(obs.st, oracle.config_policy(1e-10, 1e-12),
                           oracle.events(1000.0, horizon_offset=1e-6), y, -1.0)
        assert out.termination in (oracle.TERM_ESCAPE, oracle.TERM_HORIZON)
        assert out.drift_energy == 0.0
        assert max(out.drift_lz, out.drift_carter, out.drift_null) <= 1e-7


def test_bpt_isco_and_special_cases(oracle):
    def bpt(a):
        z1 = 1 + (1 - a * a) ** (1 / 3) * ((1 + a) ** (1 / 3) + (1 - a) ** (1 / 3))
        z2 = math.sqrt(3 * a * a + z1 * z1)
        return 3 + z2 - math.sqrt((3 - z1) * (3 + z1 + 2 * z2))
    for a in (0.0, 0.5, 0.8, 0.94, 0.998):
        assert abs(oracle.isco_prograde(a, 0.0) - bpt(a)) <= 1e-12
    assert abs(oracle.isco_prograde(0.0, 1.0) - 4.0) <= 1e-12  # extremal RN
    assert abs(oracle.isco_prograde(0.94, 0.0) - 2.0235931047) <= 1e-9
    assert abs(oracle.isco_prograde(0.8, 0.0) - 2.9066438545) <= 1e-9


def synge_shadow_angle(charge: float, observer_radius: float) -> float:
    """Angular radius of the shadow seen by a static observer in Reissner-
    Nordstrom (M = 1), Synge's formula: sin(alpha) = b_c sqrt(f(r_o)) / r_o with
    f(r) = 1 - 2/r + Q^2/r^2, photon sphere r_ph = (3 + sqrt(9 - 8 Q^2)) / 2 and
    critical impact parameter b_c = r_ph / sqrt(f(r_ph))."""
    q2 = charge * charge
    rph = 0.5 * (3.0 + math.sqrt(9.0 - 8.0 * q2))
    f = lambda r: 1.0 - 2.0 / r + q2 / (r * r)
    return math.asin(rph / math.sqrt(f(rph)) * math.sqrt(f(observer_radius)) / observer_radius)


@pytest.mark.parametrize("charge,chart,radius,inclination", [
    (0.0, 1.0, 30.0, 60.0), (0.8, -1.0, 30.0, 35.0), (0.99, 1.0, 12.0, 80.0)])
def test_static_observer_shadow_edge_is_synges(oracle, charge, chart, radius, inclination):
    """A pinhole view from a static observer at a = 0 (the scene camera of
    scene.rs:211-271): rays at 1e-6 relative inside Synge's angle, in six
    directions on the sight plane, fall into the horizon and rays 1e-6
    outside escape, in both charts. A wrong Q^2 sign or a dropped lapse
    factor moves the edge by more than 1e-2."""
    O = oracle
    alpha = synge_shadow_angle(charge, radius)
    xyz = O.oblate_to_cartesian(0.0, chart, radius, math.radians(inclination), 0.0)
    obs = O.observation(1.0, 0.0, charge, chart, [0.0, *xyz], [0, 0, 0, 0], [0, 0, 1], 1.0, 64, 64,
                        math.radians(40.0))
    ev = O.events(1000.0, None, horizon_offset=1e-6)
    eps = np.array([-1e-3, -1e-6, 1e-6, 1e-3])
    psi = np.linspace(0.0, 2.0 * math.pi, 7)[:-1]
    E, P = np.meshgrid(eps, psi)
    t = np.tan(alpha * (1.0 + E.ravel()))
    sxy = np.stack([t * np.cos(P.ravel()), t * np.sin(P.ravel())], axis=-1)
    out, st = O.observation_trace_ordered(obs, O.config_policy(1e-12, 1e-14), ev, sxy, O.Emitter(), 8)
    assert np.all(st == 0)
    term = O.outcomes_to_arrays(out)["termination"].reshape(E.shape)
    assert np.all(term[:, :2] == O.TERM_HORIZON)
    assert np.all(term[:, 2:] == O.TERM_ESCAPE)
