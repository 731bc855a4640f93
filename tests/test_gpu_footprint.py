"""Source-chart footprint on the GPU against the oracle (needs an MI355X).

The reference's gate (docs/validation.md, surface-footprint section): the
centre and its four real +-0.25-pixel neighbours must all end unambiguously
on the surface with one branch key, and the GPU and CPU Jacobians of
(r, r_c dphi) against the pixel coordinates agree to 3e-3 in the matrix
max-norm relative error (crates/gravlume-render/src/gpu_trace_tests/
surface.rs:70-100). Here on the fixture pixel (640, 16) of the default
a = 0.8 view with the surface [6, 20] M, and on 64 seeded surface pixels.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GATE = 3e-3


@pytest.fixture(scope="module")
def gl():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gravlume_amd
    return gravlume_amd


def test_surface_footprint_matches_the_oracle(gl, oracle):
    from gravlume_amd import scene as S
    import torch
    sc = S.Scene(spin=0.8, charge=0.0, inclination_deg=60.0, observer_radius=30.0, width=1280,
                 height=720, fov_deg=45.0, chart=S.INGOING, disk=(6.0, 20.0), escape_radius=200.0,
                 escape_map=False, emission="bolometric")
    pipe = gl.TracePipeline([sc])
    tgt = pipe.create_target(0)
    pipe.encode_batch(tgt)
    torch.cuda.synchronize()
    h = tgt.to_host()
    srf = np.flatnonzero(h["termination"] == 7)
    pix = np.concatenate([[16 * 1280 + 640], np.random.default_rng(9).choice(srf, 64, replace=False)])
    gpu = pipe.surface_footprint(0, pix)
    obs = oracle.observation(1.0, sc.spin, sc.charge, float(sc.chart), [0.0, *sc.observer_xyz()],
                             [0, 0, 0, 0], [0, 0, 1], 1.0, sc.width, sc.height,
                             math.radians(sc.fov_deg))
    ev = oracle.events(sc.escape_radius, sc.disk_radii(), horizon_offset=sc.horizon_offset)
    cpu = oracle.surface_footprint(obs, oracle.config_policy(1e-10, 1e-12), ev, oracle.Emitter(),
                                   pix % sc.width, pix // sc.width)
    both = gpu["resolved"] & (cpu["status"] == 0)
    print("resolved gpu", gpu["resolved"].sum(), "cpu", (cpu["status"] == 0).sum(), "both", both.sum())
    assert gpu["resolved"][0] and cpu["status"][0] == 0
    assert np.mean(gpu["resolved"] == (cpu["status"] == 0)) >= 0.9
    assert both.sum() >= 40
    err = np.max(np.abs(gpu["J"] - cpu["J"]), axis=(1, 2)) / np.max(np.abs(cpu["J"]), axis=(1, 2))
    print("max-norm relative error: fixture", err[0], "worst", err[both].max(),
          "median", np.median(err[both]))
    assert err[0] <= GATE
    assert np.all(err[both] <= GATE)
    assert np.array_equal(gpu["parity"][both], cpu["parity"][both])
    assert np.allclose(gpu["singular"][both], cpu["singular"][both], rtol=5e-3, atol=1e-6)
    # a pixel off the surface is never resolved
    off = np.flatnonzero(h["termination"] != 7)[:8]
    assert not pipe.surface_footprint(0, off)["resolved"].any()


def test_adjoint_footprint_matches_the_oracle(gl, oracle):
    """The analytic footprint from two discrete-adjoint passes (seed r, then
    phi_s; gl_vjp_surface_footprint) against f64 central differences of the
    oracle at +-0.02 pixel and rtol 1e-12, on the fixture pixel and 48 seeded
    resolved pixels of the default view, within the reference's 3e-3
    max-norm gate; parity and singular values as the oracle's, and the
    five-ray GPU footprint on the same pixels within the same gate."""
    from gravlume_amd import scene as S
    import torch
    sc = S.Scene(spin=0.8, charge=0.0, inclination_deg=60.0, observer_radius=30.0, width=1280,
                 height=720, fov_deg=45.0, chart=S.INGOING, disk=(6.0, 20.0), escape_radius=200.0,
                 escape_map=False, emission="bolometric")
    pipe = gl.TracePipeline([sc])
    vt = gl.VjpTrace(pipe, 0)
    torch.cuda.synchronize()
    fp = vt.surface_footprint()
    h = vt.target.to_host()
    srf = h["termination"] == 7
    # resolved only on unambiguous surface hits (and on all of them that stay
    # within the adjoint's step capacity)
    unamb = srf & (h["flags"] == 0) & (h["candidates"] == 4)
    assert not (fp["resolved"] & ~unamb).any()
    assert fp["resolved"].sum() >= 0.999 * unamb.sum()
    assert not fp["J"][~fp["resolved"]].any()
    pool = np.flatnonzero(fp["resolved"])
    pix = np.concatenate([[16 * 1280 + 640], np.random.default_rng(21).choice(pool, 48, replace=False)])
    assert fp["resolved"][pix[0]]
    obs = oracle.observation(1.0, sc.spin, sc.charge, float(sc.chart), [0.0, *sc.observer_xyz()],
                             [0, 0, 0, 0], [0, 0, 1], 1.0, sc.width, sc.height,
                             math.radians(sc.fov_deg))
    ev = oracle.events(sc.escape_radius, sc.disk_radii(), horizon_offset=sc.horizon_offset)
    cpu = oracle.surface_footprint(obs, oracle.config_policy(1e-12, 1e-14), ev, oracle.Emitter(),
                                   pix % sc.width, pix // sc.width, offset=0.02)
    ok = cpu["status"] == 0
    assert ok[0] and ok.sum() >= 40
    Jv = fp["J"][pix]
    err = np.max(np.abs(Jv - cpu["J"]), axis=(1, 2)) / np.max(np.abs(cpu["J"]), axis=(1, 2))
    print("adjoint footprint max-norm relative error: fixture", err[0], "worst", err[ok].max(),
          "median", np.median(err[ok]))
    assert np.all(err[ok] <= GATE)
    assert np.array_equal(fp["parity"][pix][ok], cpu["parity"][ok])
    assert np.allclose(fp["singular"][pix][ok], cpu["singular"][ok], rtol=5e-3, atol=1e-6)
    fd = pipe.surface_footprint(0, pix)
    both = fd["resolved"] & ok
    assert both.sum() >= 40
    efd = np.max(np.abs(fd["J"][both] - Jv[both]), axis=(1, 2)) / np.max(np.abs(Jv[both]), axis=(1, 2))
    print("five-ray GPU footprint vs adjoint: worst", efd.max())
    assert np.all(efd <= GATE)
    vt.close()
