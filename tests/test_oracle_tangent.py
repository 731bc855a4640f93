"""The oracle's forward-mode tangent integration (oracle/tangent.cpp) against
central differences (h = 1e-5) of the oracle's own f64 trace at rtol 1e-12:
dg/d(a, Q, inclination, s_x, s_y) of surface-hit pixels of the 512x512
adjoint scene (a = 0.9, Q = 0.2, inclination 70 deg). Both are independent
f64 routes to the same derivative; agreement to 1e-5 relative pins the
tangent integration that the GPU adjoint is checked against."""
import math

import numpy as np


def test_tangent_matches_central_differences(oracle):
    from gravlume_amd import scene as S
    O = oracle
    sc = S.adjoint_scene()
    disk = sc.disk_radii()
    obs = O.observation(1.0, sc.spin, sc.charge, float(sc.chart), [0.0, *sc.observer_xyz()],
                        [0, 0, 0, 0], [0, 0, 1], 1.0, sc.width, sc.height, math.radians(sc.fov_deg))
    px = np.array([256, 300, 200, 380, 150, 330])
    py = np.array([300, 320, 330, 290, 280, 400])
    sxy = O.sight_planes(obs, px, py)
    g, dg, st = O.tangent_frequency_ratio(sc.spin, sc.charge, math.radians(sc.inclination_deg),
                                          float(sc.chart), sc.observer_radius, sc.width,
                                          sc.height, math.radians(sc.fov_deg), sxy, disk,
                                          sc.escape_radius, rtol=1e-10, threads=4)
    hit = st == 0
    assert hit.sum() >= 4
    ev = O.events(sc.escape_radius, disk, horizon_offset=sc.horizon_offset)

    def traced(a, q, inc, s2):
        xyz = S.oblate_to_cartesian(a, sc.chart, sc.observer_radius, inc, 0.0)
        ob = O.observation(1.0, a, q, float(sc.chart), [0.0, *xyz], [0, 0, 0, 0], [0, 0, 1], 1.0,
                           sc.width, sc.height, math.radians(sc.fov_deg))
        outs, _ = O.observation_trace_ordered(ob, O.config_policy(1e-12, 1e-14), ev, s2,
                                              O.Emitter(), 4)
        A = O.outcomes_to_arrays(outs)
        return A["termination"], A["frequency_ratio"]

    base = [sc.spin, sc.charge, math.radians(sc.inclination_deg)]
    term0, g0 = traced(*base, sxy)
    assert np.all(term0[hit] == 7)
    assert np.max(np.abs(g[hit] - g0[hit]) / g0[hit]) <= 1e-9
    h = 1e-5
    for k in range(5):
        gs = []
        for sgn in (1.0, -1.0):
            p, s2 = list(base), sxy.copy()
            if k < 3:
                p[k] += sgn * h
            else:
                s2[:, k - 3] += sgn * h
            gs.append(traced(p[0], p[1], p[2], s2)[1])
        fd = (gs[0] - gs[1]) / (2 * h)
        rel = np.abs(dg[hit, k] - fd[hit]) / np.abs(fd[hit])
        assert np.all(rel <= 1e-5), (k, rel)
