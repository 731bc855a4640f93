"""The oracle's source-chart footprint (oracle.surface_footprint, after
footprint.rs:74-151) against what the symmetry of a face-on Schwarzschild view
fixes: with the observer on the disk's axis (r = 30 M, a = 0, outgoing chart,
80 deg FOV, 200x200, disk [6, 20] M), the source radius depends only on the
pixel's distance rho from the image centre and the source azimuth turns with
the pixel's position angle. So, at every resolved pixel,

* the r row (dr/dx, dr/dy) is parallel to the pixel's offset from the centre,
* the r_c dphi row is perpendicular to it, with length r_c / rho (the
  azimuth changes by 1 / rho per pixel across the offset),
* and every direct-image pixel has the same parity.

A transposed Jacobian, a missing r_c factor or a Jacobian per sight-plane
unit rather than per pixel fails these by O(1); the central differences at
+-0.25 pixel meet them to ~2e-5.
"""
import math

import numpy as np


def test_face_on_schwarzschild_footprint_invariants(oracle):
    W = H = 200
    obs = oracle.observation(1.0, 0.0, 0.0, -1.0, [0.0, 0.0, 0.0, 30.0], [0, 0, 0, 0],
                             [1.0, 0.0, 0.0], 1.0, W, H, math.radians(80.0))
    ev = oracle.events(200.0, (6.0, 20.0), horizon_offset=1e-6)
    pol = oracle.config_policy(1e-10, 1e-12)
    rng = np.random.default_rng(0)
    px, py = rng.integers(0, W, 200), rng.integers(0, H, 200)
    fp = oracle.surface_footprint(obs, pol, ev, oracle.Emitter(), px, py, threads=8)
    res = fp["status"] == 0
    assert res.sum() >= 80
    outs, st = oracle.observation_trace_ordered(obs, pol, ev, oracle.sight_planes(obs, px, py),
                                                oracle.Emitter(), 8)
    rc = oracle.outcomes_to_arrays(outs)["source_radius"][res]
    J = fp["J"][res]
    vx, vy = (px + 0.5 - W / 2.0)[res], (py + 0.5 - H / 2.0)[res]
    rho = np.hypot(vx, vy)
    n0, n1 = np.linalg.norm(J[:, 0, :], axis=1), np.linalg.norm(J[:, 1, :], axis=1)
    parallel = (J[:, 0, 0] * vy - J[:, 0, 1] * vx) / (n0 * rho)
    perpendicular = (J[:, 1, 0] * vx + J[:, 1, 1] * vy) / (n1 * rho)
    assert np.abs(parallel).max() < 2e-4
    assert np.abs(perpendicular).max() < 2e-4
    assert np.abs(n1 * rho / rc - 1.0).max() < 2e-4
    assert len(np.unique(fp["parity"][res])) == 1
