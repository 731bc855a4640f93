"""Image-space accelerators on the GPU, through the C ABI (needs an MI355X).

* Accelerated sky trace (geodesic_acceleration.wgsl:330-601) on the four
  view profiles of the reference's transfer test
  (crates/gravlume-render/src/gpu_trace_tests.rs:630-690): against the full
  trace of the same scene, every pixel keeps its termination class and event
  candidates, traced pixels are bit-identical, and reconstructed escape
  directions lie within the reference's 3.82e-4 chord budget.
* Reconstructed pixels against the f64 oracle: on the default profile, the
  blended escape direction of 512 seeded reconstructed pixels lies within
  the same 3.82e-4 chord of the oracle's traced direction (rtol 1e-10).
* Shadow-edge coverage (shadow_coverage.wgsl:1-136): the edge list is stable
  under repetition, only edge texels change, and the coverage of sampled edge
  pixels equals the f64 oracle's count of escaping sub-samples wherever the
  oracle's classification of the sub-samples is settled.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHORD_BUDGET = 3.82e-4
SUBSAMPLES = [(0.375, 0.125), (0.875, 0.375), (0.625, 0.875), (0.125, 0.625)]


@pytest.fixture(scope="module")
def gl():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gravlume_amd
    return gravlume_amd


PROFILES = [("default", 0.8, 0.0, 30.0, math.pi / 4), ("near-extreme-wide", 0.99, 0.0, 30.0, math.pi / 2),
            ("negative-near", -0.8, 0.0, 5.0, math.pi / 4), ("kerr-newman", 0.6, 0.5, 30.0, math.pi / 4)]


def _profile(S, spin, charge, r_obs, fov, **kw):
    """Outgoing chart, static observer at oblate (r_obs, 60 deg, 0) looking at
    the origin, 1280x720, escape at 200 M, no surface (gpu_trace_tests.rs:788-830)."""
    base = dict(spin=spin, charge=charge, inclination_deg=60.0, observer_radius=r_obs, width=1280,
                height=720, fov_deg=math.degrees(fov), chart=S.OUTGOING, disk=None,
                escape_radius=200.0, capture_certificate=abs(spin) <= 0.9 and r_obs <= 256)
    base.update(kw)
    return S.Scene(**base)


@pytest.mark.parametrize("profile", PROFILES, ids=[p[0] for p in PROFILES])
def test_accelerated_sky_preserves_the_full_trace(gl, profile):
    from gravlume_amd import scene as S
    import torch
    label, spin, charge, r_obs, fov = profile
    sc = _profile(S, spin, charge, r_obs, fov)
    pipe = gl.TracePipeline([sc])
    full, acc = pipe.create_target(0), pipe.create_target(0)
    pipe.encode_batch(full)
    pipe.encode_accelerated(acc)
    torch.cuda.synchronize()
    hf, ha = full.to_host(), acc.to_host()
    assert np.array_equal(ha["termination"], hf["termination"]), label
    assert np.array_equal(ha["candidates"], hf["candidates"]), label
    rec = (ha["flags"] & 128) != 0
    print(label, "reconstructed fraction", rec.mean())
    assert np.all(ha["termination"][rec] == 2)
    traced = ~rec
    for key in ("observable", "travel_time", "affine", "steps", "rejected", "flags",
                "radial_turnings", "equatorial_crossings", "azimuth_winding", "drift"):
        assert np.array_equal(ha[key][traced], hf[key][traced]), (label, key)
    assert np.array_equal(ha["rgba16f"][traced].view(np.uint16), hf["rgba16f"][traced].view(np.uint16))
    esc = hf["termination"] == 2
    chord = np.linalg.norm(ha["observable"][esc].astype(np.float64) - hf["observable"][esc], axis=1)
    print(label, "max chord", chord.max())
    assert chord.max() <= CHORD_BUDGET, (label, chord.max())
    if r_obs >= 30.0:
        assert rec.sum() > 0
    # reconstructed pixels: unit directions, escape tag, zero traced-only fields
    assert np.all(np.abs(np.linalg.norm(ha["observable"][rec], axis=1) - 1) < 1e-5)
    assert np.all(ha["rgba16f"][rec][:, 3] == 1.0)
    assert np.all(ha["steps"][rec] == 0) and np.all(ha["travel_time"][rec] == 0)


def test_accelerated_sky_rejects_a_surface_scene(gl):
    from gravlume_amd import scene as S
    from gravlume_amd import _capi
    sc = _profile(S, 0.8, 0.0, 30.0, math.pi / 4, disk=(6.0, 20.0), width=64, height=36,
                  capture_certificate=False)
    pipe = gl.TracePipeline([sc])
    with pytest.raises(_capi.GravlumeError):
        pipe.encode_accelerated(pipe.create_target(0))


def _texel(rgb, d, margin=1e-3):
    _, H, W = rgb.shape
    lon = np.arctan2(d[:, 1], d[:, 0])
    colat = np.arccos(np.clip(d[:, 2], -1.0, 1.0))
    fu, fv = (lon + np.pi) * W / (2 * np.pi), colat * H / np.pi
    iu = np.clip(fu.astype(np.int64), 0, W - 1)
    iv = np.clip(fv.astype(np.int64), 0, H - 1)
    clear = (np.abs(fu - np.round(fu)) > margin) & (np.abs(fv - np.round(fv)) > margin)
    return rgb[:, iv, iu].T, clear


def test_shadow_refinement_coverage_matches_the_oracle(gl, oracle):
    """The reference's default view (a = 0.8, outgoing chart, r = 30 M, 60 deg,
    FOV 45 deg; gpu_trace_tests.rs:849-855) at 320x180 with the escape map."""
    from gravlume_amd import scene as S
    import torch
    sc = S.Scene(spin=0.8, charge=0.0, inclination_deg=60.0, observer_radius=30.0, width=320,
                 height=180, fov_deg=45.0, chart=S.OUTGOING, disk=None, escape_radius=200.0)
    pipe = gl.TracePipeline([sc])
    tgt = pipe.create_target(0)
    pipe.encode_batch(tgt)
    torch.cuda.synchronize()
    before = tgt.to_host()
    pipe.refine_shadow(tgt)
    n1 = pipe.shadow_edges()
    once = tgt.to_host()
    pipe.refine_shadow(tgt)
    n2 = pipe.shadow_edges()
    twice = tgt.to_host()
    assert n1 > 0 and n1 == n2
    assert np.array_equal(once["rgba16f"].view(np.uint16), twice["rgba16f"].view(np.uint16))
    # the edge set, recomputed on the host from the status plane
    t = before["termination"].reshape(sc.height, sc.width)
    cls = np.where(t == 1, 0, np.where(t == 2, 1, -1))
    edge = np.zeros_like(cls, dtype=bool)
    for dy, dx in ((0, 1), (0, -1), (1, 0), (-1, 0)):
        nb = np.full_like(cls, -1)
        ys = slice(max(dy, 0), sc.height + min(dy, 0))
        yd = slice(max(-dy, 0), sc.height + min(-dy, 0))
        xs = slice(max(dx, 0), sc.width + min(dx, 0))
        xd = slice(max(-dx, 0), sc.width + min(-dx, 0))
        nb[yd, xd] = cls[ys, xs]
        edge |= (cls >= 0) & (nb == 1 - cls)
    edge = edge.ravel()
    assert edge.sum() == n1
    changed = np.any(once["rgba16f"].view(np.uint16) != before["rgba16f"].view(np.uint16), axis=1)
    assert not np.any(changed & ~edge)
    alpha = once["rgba16f"][edge, 3].astype(np.float64)
    assert np.all(np.isin(alpha, [0.0, 0.25, 0.5, 0.75, 1.0]))
    # oracle sub-samples of every edge pixel
    idx = np.flatnonzero(edge)
    obs = oracle.observation(1.0, sc.spin, sc.charge, float(sc.chart), [0.0, *sc.observer_xyz()],
                             [0, 0, 0, 0], [0, 0, 1], 1.0, sc.width, sc.height,
                             math.radians(sc.fov_deg))
    ev = oracle.events(sc.escape_radius, None, horizon_offset=sc.horizon_offset)
    sxy = np.concatenate([oracle.sight_planes(obs, idx % sc.width, idx // sc.width, ox, oy)
                          for ox, oy in SUBSAMPLES]).reshape(4, -1, 2).transpose(1, 0, 2).reshape(-1, 2)
    ref, st = oracle.observation_trace_ordered(obs, oracle.config_policy(1e-10, 1e-12), ev, sxy, None, 8)
    fine, _ = oracle.observation_trace_ordered(obs, oracle.config_policy(1e-12, 1e-14), ev, sxy, None, 8)
    A, F = oracle.outcomes_to_arrays(ref), oracle.outcomes_to_arrays(fine)
    term = A["termination"].reshape(-1, 4)
    settled = np.all(term == F["termination"].reshape(-1, 4), axis=1) & np.all(np.isin(term, [1, 2]), axis=1)
    esc_count = (term == 2).sum(axis=1)
    got = once["rgba16f"][idx, 3].astype(np.float64) * 4
    agree = got[settled] == esc_count[settled]
    print("edge pixels", n1, "sampled", idx.size, "settled", settled.sum(), "agree", agree.mean())
    assert settled.sum() >= 0.6 * idx.size
    assert agree.mean() >= 0.97
    # colour: the mean escape-map texel of the escaping sub-samples, checked on
    # pixels whose escaping sub-samples are well conditioned (the oracle's
    # direction moves by < 3e-5 rad for a 1e-7 shift of the sight point, the
    # size of the f32 rounding of the initial direction) and lie at least
    # 0.05 texel from a texel edge; near the critical curve the direction of
    # an escaping ray amplifies any difference and no f32/f64 match is owed
    shifted = sxy + np.array([1e-7, 0.0])
    moved, _ = oracle.observation_trace_ordered(obs, oracle.config_policy(1e-10, 1e-12), ev, shifted,
                                                None, 8)
    M = oracle.outcomes_to_arrays(moved)
    dirs = A["escape_direction"].reshape(-1, 4, 3)
    sens = np.linalg.norm(M["escape_direction"].reshape(-1, 4, 3) - dirs, axis=2)
    rgb = pipe.escape_map_rgb
    ok = 0
    for k in np.flatnonzero(settled)[agree]:
        e = term[k] == 2
        if not e.any() or np.any(sens[k][e] >= 3e-5):
            continue
        tex, clear = _texel(rgb, dirs[k][e], margin=0.05)
        if not clear.all():
            continue
        want = tex.sum(axis=0) / 4
        have = once["rgba16f"][idx[k], :3].astype(np.float64)
        assert np.all(np.abs(have - want) <= 1e-3), (k, have, want)
        ok += 1
    print("colour-checked edge pixels", ok)
    assert ok >= 10


def test_reconstructed_pixels_match_the_oracle(gl, oracle):
    from gravlume_amd import scene as S
    import torch
    sc = _profile(S, 0.8, 0.0, 30.0, math.pi / 4)
    pipe = gl.TracePipeline([sc])
    acc = pipe.create_target(0)
    pipe.encode_accelerated(acc)
    torch.cuda.synchronize()
    h = acc.to_host()
    rec = np.flatnonzero((h["flags"] & 128) != 0)
    assert rec.size > 10000
    idx = np.sort(np.random.default_rng(23).choice(rec, 512, replace=False))
    obs = oracle.observation(1.0, sc.spin, sc.charge, float(sc.chart), [0.0, *sc.observer_xyz()],
                             [0, 0, 0, 0], [0, 0, 1], 1.0, sc.width, sc.height,
                             math.radians(sc.fov_deg))
    ev = oracle.events(sc.escape_radius, None, horizon_offset=sc.horizon_offset)
    sxy = oracle.sight_planes(obs, idx % sc.width, idx // sc.width)
    outs, st = oracle.observation_trace_ordered(obs, oracle.config_policy(1e-10, 1e-12), ev, sxy,
                                                None, 16)
    A = oracle.outcomes_to_arrays(outs)
    assert np.all(A["termination"] == 2)
    chord = np.linalg.norm(h["observable"][idx].astype(np.float64) - A["escape_direction"], axis=1)
    print("reconstructed vs oracle chord: max", chord.max(), "p99", np.quantile(chord, 0.99))
    assert chord.max() <= CHORD_BUDGET
