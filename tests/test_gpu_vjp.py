"""Discrete-adjoint VJP of the per-pixel frequency ratio g against f64 central
differences of the oracle and against the oracle's forward-mode tangent
integration (oracle/tangent.cpp) (needs an MI355X).

Workload: 512x512, a = 0.9, Q = 0.2, inclination 70 deg, r_obs = 100 M, FOV
20 deg, thin disk [r_ISCO, 25 M], f32 forward at rtol 1e-6, checkpoints every
32 accepted steps. 1,024 seeded surface-hit pixels (away from the disk
edges) are differentiated by the oracle at rtol 1e-12 with h = 1e-5 in a, Q,
inclination and the sight-plane coordinates (s_x, s_y); pixels whose class
changes under any perturbation are dropped.

Tolerances: the aggregated parameter gradient sum_i w_i dg_i/dp (positive
seeded weights) within 1e-4 relative of the differences; per-pixel direction
gradients within 1e-4 relative at the median and 1e-3 at the 90th
percentile (measured on MI355X: 4.8e-6, 2.6e-6 and 1.9e-5).
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


def _oracle_g(O, sc, a, q, incl_deg, sxy, disk, threads=16):
    from gravlume_amd import scene as S
    xyz = S.oblate_to_cartesian(a, sc.chart, sc.observer_radius, math.radians(incl_deg), 0.0)
    obs = O.observation(1.0, a, q, float(sc.chart), [0.0, *xyz], [0, 0, 0, 0], [0, 0, 1], 1.0,
                        sc.width, sc.height, math.radians(sc.fov_deg))
    ev = O.events(sc.escape_radius, disk, horizon_offset=sc.horizon_offset)
    outs, st = O.observation_trace_ordered(obs, O.config_policy(1e-12, 1e-14), ev, sxy,
                                           O.Emitter(), threads)
    assert np.all(st == 0)
    A = O.outcomes_to_arrays(outs)
    return A["termination"], A["frequency_ratio"]


def test_vjp_matches_oracle_central_differences(gl, oracle):
    import torch
    from gravlume_amd import scene as S
    sc = S.adjoint_scene()
    pipe = gl.TracePipeline([sc])
    vjp = gl.VjpTrace(pipe, 0, checkpoint_interval=32)
    torch.cuda.synchronize()
    h = vjp.target.to_host()
    disk = sc.disk_radii()
    r = h["observable"][:, 0]
    ok = (h["termination"] == 7) & (r > disk[0] * 1.05) & (r < disk[1] * 0.95)
    rng = np.random.default_rng(20261207)
    idx = np.sort(rng.choice(np.flatnonzero(ok), 1024, replace=False))
    obs = oracle.observation(1.0, sc.spin, sc.charge, float(sc.chart), [0.0, *sc.observer_xyz()],
                             [0, 0, 0, 0], [0, 0, 1], 1.0, sc.width, sc.height,
                             math.radians(sc.fov_deg))
    sxy = oracle.sight_planes(obs, idx % sc.width, idx // sc.width)
    eps = 1e-5
    base = (sc.spin, sc.charge, sc.inclination_deg)
    fd = np.zeros((idx.size, 5))
    valid = np.ones(idx.size, dtype=bool)
    for k in range(5):
        gs = []
        for sgn in (1.0, -1.0):
            a, q, inc = base
            s2 = sxy.copy()
            if k == 0:
                a += sgn * eps
            elif k == 1:
                q += sgn * eps
            elif k == 2:
                inc = math.degrees(math.radians(inc) + sgn * eps)
            else:
                s2[:, k - 3] += sgn * eps
            term, g = _oracle_g(oracle, sc, a, q, inc, s2, disk)
            valid &= term == 7
            gs.append(g)
        fd[:, k] = (gs[0] - gs[1]) / (2 * eps)
    assert valid.mean() > 0.9
    idx, fd = idx[valid], fd[valid]
    w = rng.uniform(0.5, 1.5, idx.size)
    grad_g = torch.zeros(sc.rays, dtype=torch.float32, device="cuda")
    grad_g[torch.from_numpy(idx).cuda()] = torch.from_numpy(w.astype(np.float32)).cuda()
    gp, gdir = vjp.backward(grad_g, direction=True)
    want = (w[:, None] * fd[:, :3]).sum(axis=0)
    rel = np.abs(gp - want) / np.abs(want)
    print("params", gp, want, rel)
    assert np.all(rel <= 1e-4), (gp, want, rel)
    gd = gdir.cpu().numpy()[:, idx].T / w[:, None]
    rel_d = np.abs(gd - fd[:, 3:]) / np.maximum(np.abs(fd[:, 3:]), 1e-3 * np.abs(fd[:, 3:]).max())
    print("direction rel: median", np.median(rel_d), "p90", np.quantile(rel_d, 0.9))
    assert np.median(rel_d) <= 1e-4
    assert np.quantile(rel_d, 0.9) <= 1e-3
    # the same VJP against the oracle's forward-mode tangent integration
    tg_g, tg_dg, tg_st = oracle.tangent_frequency_ratio(
        sc.spin, sc.charge, math.radians(sc.inclination_deg), float(sc.chart),
        sc.observer_radius, sc.width, sc.height, math.radians(sc.fov_deg), sxy[valid], disk,
        sc.escape_radius, rtol=1e-10, threads=16)
    ok_t = tg_st == 0
    assert ok_t.mean() > 0.95
    want_t = (w[ok_t, None] * tg_dg[ok_t, :3]).sum(axis=0)
    got_t = vjp.backward(torch.where(
        torch.isin(torch.arange(sc.rays, device="cuda"), torch.from_numpy(idx[ok_t]).cuda()),
        grad_g, torch.zeros_like(grad_g)))
    rel_t = np.abs(got_t - want_t) / np.abs(want_t)
    print("params vs tangents", got_t, want_t, rel_t)
    assert np.all(rel_t <= 1e-4), (got_t, want_t, rel_t)
    rel_dt = np.abs(gd[ok_t] - tg_dg[ok_t, 3:]) / np.maximum(np.abs(tg_dg[ok_t, 3:]),
                                                            1e-3 * np.abs(tg_dg[ok_t, 3:]).max())
    print("direction vs tangents: median", np.median(rel_dt), "p90", np.quantile(rel_dt, 0.9))
    assert np.median(rel_dt) <= 1e-4
    assert np.quantile(rel_dt, 0.9) <= 1e-3
    # pixels off the surface or with zero weight get no gradient
    off = np.setdiff1d(np.arange(sc.rays), idx)
    assert np.all(gdir.cpu().numpy()[:, off] == 0)
    vjp.close()


def test_autograd_function_contract(gl):
    """FrequencyRatioImage follows torch.autograd.Function: the loss
    sum (g - g*)^2 over surface pixels back-propagates to (a, Q, inclination)."""
    import torch
    from gravlume_amd import scene as S
    sc = S.adjoint_scene(width=128, height=128)
    params = torch.tensor([0.9, 0.2, math.radians(70.0)], dtype=torch.float64, requires_grad=True)
    g = gl.FrequencyRatioImage.apply(params, sc)
    with torch.no_grad():
        target = gl.FrequencyRatioImage.apply(
            torch.tensor([0.92, 0.18, math.radians(71.0)], dtype=torch.float64), sc)
    mask = (g.detach() > 0) & (target > 0)
    loss = ((g - target)[mask] ** 2).sum()
    loss.backward()
    assert params.grad is not None and torch.all(torch.isfinite(params.grad))
    assert torch.any(params.grad != 0)


def test_vjp_forward_is_the_trace_where_it_differentiates(gl):
    """The adjoint's recording forward gives the plain trace's outputs bit for
    bit on every surface pixel it differentiates; the surface pixels it
    cannot differentiate (more accepted steps than it records, or a visit
    to the f64 near-horizon island, which it does not take) carry
    GL_STATUS_NO_ADJOINT and receive no direction gradient."""
    import torch
    from gravlume_amd import scene as S
    sc = S.adjoint_scene(width=256, height=256)
    pipe = gl.TracePipeline([sc])
    plain = pipe.create_target(0)
    pipe.encode_batch(plain)
    vjp = gl.VjpTrace(pipe, 0, checkpoint_interval=32)
    torch.cuda.synchronize()
    hp, hv = plain.to_host(), vjp.target.to_host()
    srf = (hv["termination"] == 7) & ~hv["no_adjoint"]
    assert srf.sum() > 10000
    for k in hp:
        if k == "no_adjoint":
            continue
        assert np.array_equal(np.asarray(hp[k])[srf], np.asarray(hv[k])[srf]), k
    seed = torch.zeros(sc.rays, dtype=torch.float32, device="cuda")
    seed[torch.from_numpy(np.flatnonzero(hv["termination"] == 7)).cuda()] = 1.0
    gp, gd = vjp.backward(seed, direction=True)  # gd: (2, rays)
    nog = torch.from_numpy(hv["no_adjoint"]).cuda()
    assert torch.all(gd[:, nog] == 0)
    assert torch.any(gd[:, torch.from_numpy(srf).cuda()] != 0)
    print("surface pixels without an adjoint:", int(hv["no_adjoint"].sum()))
