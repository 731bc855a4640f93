"""HIP trace against the f64 oracle, through the C ABI (needs an MI355X).

Tolerances of the f32 DP5(4) trace at rtol 1e-6 / atol 1e-8 against the f64
oracle at rtol 1e-10 (the reference's GPU budget table, docs/validation.md
section 5.3, is the model; the numbers below are this build's). On every ray
the oracle does not flag near-critical (its class or candidates change under
a 16x tighter tolerance):
  * termination class and event id: exact on every ray with a definite GPU
    class; a ray the trace cannot settle is Uncertain (class 6), at most
    0.2% of a sample;
  * branch key (radial turnings, equatorial crossings, azimuth winding,
    polar side): exact on definite rays whose branch is settled (no
    GL_STATUS_BRANCH_UNSETTLED);
  * event residual (the event function at the localized terminal): at most
    5e-3, the reference GPU gate (gpu_trace_tests.rs:602-605);
  * escape direction: 2e-4 rad at the 99th percentile;
  * travel time: 1e-4 relative at the 99th percentile;
  * frequency ratio g: 1e-4 relative at the 99th percentile;
  * source radius: 1e-3 M at the 99th percentile;
  * null-constraint drift: 1e-4 at the 99th percentile.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = dict(uncertain=2e-3, esc_angle_p99=2e-4, tt_rel_p99=1e-4, g_rel_p99=1e-4, r_abs_p99=1e-3,
           null_p99=1e-4, residual=5e-3)


@pytest.fixture(scope="module")
def gl():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gravlume_amd
    return gravlume_amd


def _oracle_outcomes(O, sc, idx, rtol=1e-10, atol=1e-12, threads=8):
    obs = O.observation(1.0, sc.spin, sc.charge, float(sc.chart), [0.0, *sc.observer_xyz()],
                        [0, 0, 0, 0], [0, 0, 1], 1.0, sc.width, sc.height, math.radians(sc.fov_deg))
    ev = O.events(sc.escape_radius, sc.disk_radii(), horizon_offset=sc.horizon_offset)
    sxy = O.sight_planes(obs, idx % sc.width, idx // sc.width)
    em = O.Emitter()
    ref, st = O.observation_trace_ordered(obs, O.config_policy(rtol, atol), ev, sxy, em, threads)
    fine, _ = O.observation_trace_ordered(obs, O.config_policy(rtol / 16, atol / 16), ev, sxy, em,
                                          threads)
    assert np.all(st == 0)
    A = O.outcomes_to_arrays(ref)
    F = O.outcomes_to_arrays(fine)
    A["flagged"] = (A["termination"] != F["termination"]) | (A["candidates"] != F["candidates"])
    return A


def _angle(a, b):
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=-1), np.sum(a * b, axis=-1))


def _compare(h, A, idx, tol=TOL):
    ok = ~A["flagged"]
    gt, gc = h["termination"][idx], h["candidates"][idx]
    definite = ok & (gt != 6)
    unc = float(np.mean(gt[ok] == 6))
    assert unc <= tol["uncertain"], f"uncertain share {unc}"
    bad = definite & (gt != A["termination"])
    assert not bad.any(), ("class", idx[bad], gt[bad], A["termination"][bad])
    bad = definite & (gc != A["candidates"])
    assert not bad.any(), ("event id", idx[bad], gc[bad], A["candidates"][bad])
    ev = definite & np.isin(gt, (1, 2, 3, 7))
    assert np.all(np.abs(h["event_residual"][idx][ev]) <= tol["residual"])
    esc = definite & (gt == 2)
    if esc.sum() >= 20:
        a = _angle(h["observable"][idx][esc], A["escape_direction"][esc])
        assert np.quantile(a, 0.99) <= tol["esc_angle_p99"], np.quantile(a, 0.99)
        tt = np.abs(h["travel_time"][idx][esc] / A["travel_time"][esc] - 1)
        assert np.quantile(tt, 0.99) <= tol["tt_rel_p99"], np.quantile(tt, 0.99)
    srf = definite & (gt == 7)
    if srf.sum() >= 20:
        o = h["observable"][idx][srf]
        g = np.abs(o[:, 2] / A["frequency_ratio"][srf] - 1)
        assert np.quantile(g, 0.99) <= tol["g_rel_p99"], np.quantile(g, 0.99)
        r = np.abs(o[:, 0] - A["source_radius"][srf])
        assert np.quantile(r, 0.99) <= tol["r_abs_p99"], np.quantile(r, 0.99)
        dphi = np.abs((o[:, 1] - A["source_azimuth"][srf] + np.pi) % (2 * np.pi) - np.pi)
        assert np.quantile(dphi, 0.99) <= 1e-3
    assert np.quantile(h["drift"][idx][:, 0], 0.99) <= tol["null_p99"]
    # the branch key, exact where the trace settles it
    settled = definite & ~h["branch_unsettled"][idx]
    for key in ("radial_turnings", "equatorial_crossings", "azimuth_winding", "polar_side"):
        bad = settled & (h[key][idx] != A[key])
        assert not bad.any(), (key, idx[bad], h[key][idx][bad], A[key][bad])
    return dict(uncertain=unc, definite=int(definite.sum()), settled=int(settled.sum()))


def test_small_scene_matches_oracle(gl, oracle):
    """64x64, a = 0.5, Q = 0.3, observer r = 50 M at 60 deg, FOV 20 deg, disk [r_ISCO, 20]."""
    from gravlume_amd import scene as S
    import torch
    sc = S.small_cpu_scene()
    pipe = gl.TracePipeline([sc])
    tgt = pipe.create_target(0)
    pipe.encode_batch(tgt)
    torch.cuda.synchronize()
    h = tgt.to_host()
    assert np.all(h["termination"] > 0)
    idx = np.arange(sc.rays)[::4]
    A = _oracle_outcomes(oracle, sc, idx)
    _compare(h, A, idx)


def test_kerr_1080p_frame_sample_matches_oracle(gl, oracle):
    """The full 1920x1080 a = 0.94 frame; 768 seeded pixels checked against the oracle."""
    from gravlume_amd import scene as S
    import torch
    sc = S.single_frame_1080p()
    pipe = gl.TracePipeline([sc])
    tgt = pipe.create_target(0)
    pipe.encode_persistent([tgt])
    torch.cuda.synchronize()
    h = tgt.to_host()
    # size-independent properties of the whole frame
    assert np.all(h["termination"] > 0)
    assert set(np.unique(h["termination"])) <= {1, 2, 6, 7}
    unc = h["termination"] == 6
    assert unc.mean() <= 1e-4  # near-critical share of the frame
    assert np.all(h["near_critical"][unc] | (h["flags"][unc] & (8 | 16 | 32) != 0))
    assert np.all(np.isin(h["provisional"][unc], (1, 2, 7)))
    esc = h["termination"] == 2
    n = np.linalg.norm(h["observable"][esc], axis=1)
    assert np.all(np.abs(n - 1) < 1e-5)
    srf = h["termination"] == 7
    assert np.all(h["observable"][srf][:, 2] > 0)
    r_in = sc.disk_radii()[0]
    assert np.all((h["observable"][srf][:, 0] >= r_in * (1 - 1e-5)) & (h["observable"][srf][:, 0] <= 30.0 * (1 + 1e-5)))
    idx = np.random.default_rng(11).choice(sc.rays, 768, replace=False)
    A = _oracle_outcomes(oracle, sc, idx)
    _compare(h, A, idx)
    _check_radiance(h, sc, pipe.escape_map_rgb, oracle)


def test_kerr_1080p_frame_32k_pixels_exact(gl, oracle):
    """32,768 seeded pixels of the 1920x1080 a = 0.94 frame against the
    oracle: with no mismatch among them, the class agreement on unflagged
    rays is at least 99.99% at 95% confidence (rule of three: 3 / 32,768 =
    0.009%). Class, event id and the settled branch keys exact, the
    observables within the module's tolerances."""
    from gravlume_amd import scene as S
    import torch
    sc = S.single_frame_1080p()
    pipe = gl.TracePipeline([sc])
    tgt = pipe.create_target(0)
    pipe.encode_persistent([tgt])
    torch.cuda.synchronize()
    h = tgt.to_host()
    idx = np.random.default_rng(2026).choice(sc.rays, 32768, replace=False)
    A = _oracle_outcomes(oracle, sc, idx, threads=16)
    definite = ~A["flagged"] & (h["termination"][idx] != 6)
    assert definite.sum() >= 32768 - 80
    _compare(h, A, idx)


def _check_radiance(h, sc, rgb, oracle):
    """Radiance planes from the observables: the equirectangular texel of the
    escape direction, and on the disk the oracle's transport of the traced
    (r, g) (blackbody bands at g T(r) through the band table, within the
    reference's 2e-3 table budget); the RGBA16F texels are the binary16
    roundings of the planes with the class tags."""
    esc = h["termination"] == 2
    d = h["observable"][esc].astype(np.float64)
    _, H, W = rgb.shape
    lon = np.arctan2(d[:, 1], d[:, 0])
    colat = np.arccos(np.clip(d[:, 2], -1.0, 1.0))
    fu, fv = (lon + np.pi) * W / (2 * np.pi), colat * H / np.pi
    iu = np.clip(fu.astype(np.int64), 0, W - 1)
    iv = np.clip(fv.astype(np.int64), 0, H - 1)
    # texel edges within f32 rounding of the direction may round either way
    clear = (np.abs(fu - np.round(fu)) > 1e-3) & (np.abs(fv - np.round(fv)) > 1e-3)
    want = rgb[:, iv, iu].T
    got = h["radiance"][esc]
    assert np.mean(np.all(got[clear] == want[clear], axis=1)) >= 0.9999
    srf = h["termination"] == 7
    o = h["observable"][srf].astype(np.float64)
    em = oracle.Emitter(sc.intensity_at_six, 1, sc.six_temperature())
    ref = oracle.surface_radiance(em, oracle.vacuum(), o[:, 0], o[:, 2])
    assert np.all(ref["status"] == 0)
    got = h["radiance"][srf].astype(np.float64)
    table_top = 2.0 ** 24
    bad = got[:, 0] < 0
    # the band table ends at 2^24 K: hotter observed temperatures are not representable
    assert np.all(ref["observed_temperature"][bad] >= table_top * (1 - 1e-5))
    assert np.all(ref["observed_temperature"][~bad] <= table_top * (1 + 1e-5))
    rel = np.abs(got[~bad] / ref["bands"][~bad] - 1)
    assert np.max(rel) <= 2e-3, np.max(rel)
    _check_texels(h)


def _check_texels(h):
    """RGBA16F tags (lensing_preview.wgsl:64-82, surface_transport.wgsl:10):
    surface 2, escape 1, horizon 0, a failure -termination; radiance texels
    are the round-to-nearest binary16 of the f32 planes."""
    t, px, rad = h["termination"], h["rgba16f"].astype(np.float32), h["radiance"]
    srf, esc, hor = t == 7, t == 2, t == 1
    ok_srf = srf & (rad[:, 0] >= 0)
    assert np.all(px[ok_srf, 3] == 2.0) and np.all(px[esc, 3] == 1.0)
    assert np.all(px[hor] == 0.0)
    assert np.all(px[srf & ~ok_srf, 3] == -5.0)
    other = ~(srf | esc | hor)
    assert np.all(px[other, 3] == -t[other].astype(np.float32))
    both = ok_srf | esc
    assert np.array_equal(h["rgba16f"][both, :3].view(np.uint16),
                          rad[both].astype(np.float16).view(np.uint16))


@pytest.mark.parametrize("mode", [dict(refill_threshold=16, refill_interval=8, waves_per_simd=2),
                                  dict(refill_threshold=64, refill_interval=8, waves_per_simd=2),
                                  dict(refill_threshold=64, refill_interval=8, waves_per_simd=3)],
                         ids=["lane-refill", "whole-tile", "whole-tile-3-waves"])
def test_persistent_matches_plain(gl, mode):
    from gravlume_amd import scene as S
    import torch
    sc = S.single_frame_1080p(width=480, height=270)
    pipe = gl.TracePipeline([sc])
    pipe.set_persistent_config(block_size=256, blocks_per_cu=mode["waves_per_simd"], **mode)
    a, b = pipe.create_target(0), pipe.create_target(0)
    pipe.encode_batch(a)
    pipe.encode_persistent([b])
    torch.cuda.synchronize()
    ha, hb = a.to_host(), b.to_host()
    # every dispatch runs each ray on its own attempt clock (the same
    # re-projection and drift schedule) through one step routine whose
    # multiply-adds are fused per source expression only, so whole tiles and
    # lane refill give the plain dispatch's outputs bit for bit
    for k in ha:
        assert np.array_equal(np.asarray(ha[k]), np.asarray(hb[k])), k
    stats = pipe.queue_stats()
    assert stats["items"] >= sc.rays
    assert 0 < stats["attempted_steps"] <= stats["lane_slots"]


def test_ragged_viewport_and_tile_regions(gl):
    """A viewport that is not a multiple of the 8x8 tile; two tile regions
    cover it exactly like one full region (TileRegion, trace.rs:799-857)."""
    from gravlume_amd import scene as S
    import torch
    sc = S.small_cpu_scene(width=67, height=45)
    pipe = gl.TracePipeline([sc])
    full, parts = pipe.create_target(0), pipe.create_target(0)
    pipe.encode_batch(full)
    pipe.encode_batch(parts, gl.TileRegion(0, 0, 0, 9, 3))
    pipe.encode_batch(parts, gl.TileRegion(0, 0, 3, 9, 3))
    torch.cuda.synchronize()
    hf, hp = full.to_host(), parts.to_host()
    assert np.all(hf["termination"] > 0)
    assert np.array_equal(hf["termination"], hp["termination"])
    assert np.array_equal(hf["observable"], hp["observable"])
    with pytest.raises(gl.GravlumeError):
        pipe.encode_batch(parts, gl.TileRegion(0, 5, 0, 9, 1))


def test_multi_scene_persistent_matches_single_scene_launches(gl):
    from gravlume_amd import scene as S
    import torch
    scenes = S.uneven_scenes(count=6, seed=5)
    for s in scenes:
        s.width = s.height = 64
    scenes[1].width = 67  # a ragged viewport among them
    pipe = gl.TracePipeline(scenes)
    pipe.set_persistent_config(block_size=256, blocks_per_cu=2, refill_interval=8,
                               refill_threshold=16, waves_per_simd=2)  # lane refill
    multi = [pipe.create_target(i) for i in range(len(scenes))]
    pipe.encode_persistent(multi)
    single = []
    for i in range(len(scenes)):
        t = pipe.create_target(i)
        pipe.encode_persistent([t], scenes=[i])
        single.append(t)
    pipe.set_persistent_config(block_size=256, blocks_per_cu=2, refill_interval=8,
                               refill_threshold=64, waves_per_simd=2)
    tiled = [pipe.create_target(i) for i in range(len(scenes))]
    pipe.encode_persistent(tiled)
    torch.cuda.synchronize()
    for m, s, t in zip(multi, single, tiled):
        hm, hs, ht = m.to_host(), s.to_host(), t.to_host()
        assert np.all(hm["termination"] > 0)
        for k in hs:
            # one lane-refill launch over many scenes equals one per scene
            assert np.array_equal(np.asarray(hm[k]), np.asarray(hs[k])), k
        for k in hs:
            # and whole tiles trace each ray as lane refill does
            assert np.array_equal(np.asarray(ht[k]), np.asarray(hs[k])), k


def test_uneven_scenes_match_the_oracle(gl, oracle):
    """Sixteen seed-drawn scenes of the uneven multi-scene workload at their
    drawn sizes (a up to 0.998, charge, inclination 5-89 deg, 128-1024 px
    sides), traced in one whole-tile persistent launch; 256 seeded pixels of
    each against the oracle with the tolerances above."""
    from gravlume_amd import scene as S
    import torch
    scenes = S.uneven_scenes(count=16, seed=31)
    pipe = gl.TracePipeline(scenes)
    tgts = [pipe.create_target(i) for i in range(len(scenes))]
    pipe.encode_persistent(tgts)
    torch.cuda.synchronize()
    rng = np.random.default_rng(31)
    for i, (sc, t) in enumerate(zip(scenes, tgts)):
        h = t.to_host()
        assert np.all(h["termination"] > 0), i
        idx = np.sort(rng.choice(sc.rays, 256, replace=False))
        A = _oracle_outcomes(oracle, sc, idx)
        print(i, f"a={sc.spin:.3f} Q={sc.charge:.3f} inc={sc.inclination_deg:.1f} "
              f"side={sc.width}", np.bincount(A["termination"], minlength=8)[[1, 2, 7]],
              "flagged", int(A["flagged"].sum()), flush=True)
        # a up to 0.998 with pixels on the critical curve: a few of the 256
        # stay Uncertain (near-critical), none leaves with a wrong class
        _compare(h, A, idx, dict(TOL, uncertain=1.2e-2))


@pytest.mark.parametrize("charge,chart,radius,inclination", [
    (0.0, 1, 30.0, 60.0), (0.8, -1, 30.0, 35.0), (0.99, 1, 12.0, 80.0)])
def test_static_observer_shadow_is_synges_disc(gl, charge, chart, radius, inclination):
    """Schwarzschild and Reissner-Nordstrom frames (a = 0, 512x512, FOV 40
    deg, no surface): every pixel centre more than 2e-4 relative from Synge's
    shadow angle falls into the horizon inside it and escapes outside it, in
    the plain and the whole-tile dispatch."""
    from gravlume_amd import scene as S
    from test_oracle_closed_forms import synge_shadow_angle
    import torch
    sc = S.Scene(spin=0.0, charge=charge, inclination_deg=inclination, observer_radius=radius,
                 width=512, height=512, fov_deg=40.0, chart=chart, disk=None, escape_radius=1000.0,
                 escape_map=False)
    alpha = synge_shadow_angle(charge, radius)
    pipe = gl.TracePipeline([sc])
    a, b = pipe.create_target(0), pipe.create_target(0)
    pipe.encode_batch(a)
    pipe.encode_persistent([b])
    torch.cuda.synchronize()
    t = math.tan(math.radians(20.0))
    idx = np.arange(sc.rays)
    sx = t * (2.0 * (idx % 512 + 0.5) / 512 - 1.0)
    sy = t * (1.0 - 2.0 * (idx // 512 + 0.5) / 512)
    rel = np.arctan(np.hypot(sx, sy)) / alpha - 1.0
    inside, outside = rel < -2e-4, rel > 2e-4
    assert inside.sum() > 10000 and outside.sum() > 10000
    for h in (a.to_host(), b.to_host()):
        # the class the trace reaches is Synge's; close to the edge a ray may
        # stay Uncertain (near-critical), never a wrong definite class
        assert np.all(h["class_reached"][inside] == 1)
        assert np.all(h["class_reached"][outside] == 2)
        assert np.all(np.isin(h["termination"][inside], (1, 6)))
        assert np.all(np.isin(h["termination"][outside], (2, 6)))
        far = np.abs(rel) > 2e-2
        assert np.all(h["termination"][far] != 6)
        near = np.abs(rel) <= 2e-4
        assert np.all(np.isin(h["termination"][near], (1, 2, 6)))


def test_geodesic_batch_reproduces_the_b6_fixture(gl):
    """GeodesicBatch on the GPU: the Schwarzschild b = 6 scatter
    (fixtures/v1/schwarzschild-scatter-b6.toml) escapes after one turning."""
    from gravlume_amd import scene as S
    sc = S.Scene(spin=0.0, charge=0.0, chart=1, observer_radius=40.0, width=8, height=8,
                 disk=None, escape_radius=50.0, escape_map=False, horizon_offset=0.0,
                 rtol=1e-7, atol=1e-9)
    pipe = gl.TracePipeline([sc])
    y = np.array([[0, 50, 0, 0, -1.0, -0.99277494330677424337, 0.12, 0]] * 3, dtype=np.float64)
    y[1, 0] = 1.0e6  # coordinate-time origin does not matter
    out = gl.GeodesicBatch(pipe, 0).trace_ordered(y, np.ones(3))
    assert list(out["termination"]) == [2, 2, 2]
    assert list(out["radial_turnings"]) == [1, 1, 1]
    assert abs(out["travel_time"][0] - out["travel_time"][1]) <= 1e-3 * out["travel_time"][0]
    assert gl.GeodesicBatch(pipe, 0).trace_ordered(np.zeros((0, 8)), np.zeros(0))["termination"].size == 0


def test_frame_graph_replays_match_eager(gl):
    """Graph replays of a persistent frame equal the eager frame, with eager
    work on the same plan between them: a frame, and a geodesic batch four
    times the frame's size whose f64 deferrals make the plan reallocate its
    own deferral workspace (the graph replays into its own)."""
    from gravlume_amd import critical as Cr, scene as S
    import math
    import torch
    sc = S.single_frame_1080p(width=256, height=144)
    pipe = gl.TracePipeline([sc])
    eager = pipe.create_target(0)
    pipe.encode_persistent([eager])
    torch.cuda.synchronize()
    he = eager.to_host()
    tgt = pipe.create_target(0)
    graph = pipe.frame_graph([tgt])
    n = 4 * sc.rays
    states, *_ = Cr.near_critical_states(sc.spin, sc.charge, sc.observer_radius,
                                         math.radians(sc.inclination_deg), n, seed=12,
                                         chart=sc.chart)
    d_states = torch.from_numpy(states).cuda()
    d_dirs = -torch.ones(n, dtype=torch.float32, device="cuda")
    batch_out = gl.TraceTarget(n)
    for rep in range(3):
        tgt.status.zero_()
        torch.cuda.synchronize()
        graph.launch()
        graph.stream.synchronize()
        h = tgt.to_host()
        for k in he:
            assert np.array_equal(np.asarray(h[k]), np.asarray(he[k])), (rep, k)
        # eager work between replays: a frame, then the larger batch
        pipe.encode_persistent([eager])
        gl.GeodesicBatch(pipe, 0).encode(d_states, d_dirs, batch_out)
        torch.cuda.synchronize()
    assert np.all(batch_out.to_host()["termination"] > 0)


def test_capture_certificate_on_the_default_kerr_view(gl, oracle):
    """The interval capture certificate on the reference's default view
    (a = 0.8, r = 30 M, 60 deg, FOV 45 deg, 1280x720, sky only): the
    certified fraction matches the reference's 56,937 / 921,600 px; every
    certified pixel is a horizon capture in the full f32 trace and in the f64
    oracle; every other pixel traces exactly as without the certificate."""
    from gravlume_amd import scene as S
    import torch
    kw = dict(spin=0.8, charge=0.0, inclination_deg=60.0, observer_radius=30.0, width=1280,
              height=720, fov_deg=45.0, disk=None, escape_map=False)
    on_sc, off_sc = S.Scene(capture_certificate=True, **kw), S.Scene(**kw)
    on_pipe, off_pipe = gl.TracePipeline([on_sc]), gl.TracePipeline([off_sc])
    on, off = on_pipe.create_target(0), off_pipe.create_target(0)
    on_pipe.encode_batch(on)
    off_pipe.encode_batch(off)
    torch.cuda.synchronize()
    hn, hf = on.to_host(), off.to_host()
    cert = (hn["flags"] & 64) != 0
    frac = cert.mean()
    print("certified fraction", frac, cert.sum())
    assert abs(frac - 56937 / 921600) <= 1e-3
    assert np.all(hn["termination"][cert] == 1) and np.all(hn["candidates"][cert] == 2)
    assert np.all(hn["steps"][cert] == 0)
    assert np.all(hf["termination"][cert] == 1)
    assert np.array_equal(hn["termination"][~cert], hf["termination"][~cert])
    assert np.array_equal(hn["observable"][~cert], hf["observable"][~cert])
    idx = np.random.default_rng(3).choice(np.flatnonzero(cert), 256, replace=False)
    obs = oracle.observation(1.0, on_sc.spin, on_sc.charge, float(on_sc.chart),
                             [0.0, *on_sc.observer_xyz()], [0, 0, 0, 0], [0, 0, 1], 1.0,
                             on_sc.width, on_sc.height, math.radians(on_sc.fov_deg))
    sxy = oracle.sight_planes(obs, idx % on_sc.width, idx // on_sc.width)
    outs, st = oracle.observation_trace_ordered(
        obs, oracle.config_policy(1e-10, 1e-12),
        oracle.events(on_sc.escape_radius, None, horizon_offset=on_sc.horizon_offset), sxy,
        None, 8)
    assert np.all(oracle.outcomes_to_arrays(outs)["termination"] == 1)


@pytest.mark.parametrize("mode", [dict(refill_threshold=16, waves_per_simd=2),
                                  dict(refill_threshold=64, waves_per_simd=2),
                                  dict(refill_threshold=64, waves_per_simd=3)],
                         ids=["lane-refill", "whole-tile", "whole-tile-3"])
def test_mixed_work_list_matches_separate_launches(gl, mode):
    """One persistent launch over image scenes of uneven size plus, per scene,
    10% extra rays seeded next to its shadow edge as geodesic-batch requests
    (the uneven multi-scene workload in miniature) agrees with tracing each
    image and each batch in its own plain launch. In whole-tile mode the
    launch hands out the image tiles whole and then the batch requests lane
    by lane (a second cursor): every request is traced once, bit for bit as
    in its own launch."""
    from gravlume_amd import scene as S
    import torch
    scenes = S.uneven_scenes(count=6, seed=11)
    for k, s in enumerate(scenes):
        s.width = s.height = 48 + 16 * k
    pipe = gl.TracePipeline(scenes)
    pipe.set_persistent_config(block_size=256, blocks_per_cu=2, refill_interval=8, **mode)
    items, images, batches, keep = [], [], [], []
    for i, s in enumerate(scenes):
        img = pipe.create_target(i)
        items.append(("image", i, img))
        images.append(img)
        states, delta = S.shadow_edge_requests(s, 0.1, seed=100 + i)
        d_states = torch.from_numpy(states).cuda()
        d_dirs = -torch.ones(len(states), dtype=torch.float32, device="cuda")
        bt = gl.TraceTarget(len(states))
        items.append(("batch", i, d_states, d_dirs, bt))
        batches.append((bt, states, delta))
        keep.append((d_states, d_dirs))
    pipe.encode_work(items)
    torch.cuda.synchronize()
    stats = pipe.queue_stats()
    assert 0 < stats["attempted_steps"] <= stats["lane_slots"]
    for i, s in enumerate(scenes):
        ref = pipe.create_target(i)
        pipe.encode_batch(ref)
        torch.cuda.synchronize()
        hi, hr = images[i].to_host(), ref.to_host()
        assert np.all(hi["termination"] > 0)
        for k in hr:
            assert np.array_equal(np.asarray(hi[k]), np.asarray(hr[k])), (i, k)
        bt, states, delta = batches[i]
        hb = bt.to_host()
        tb = gl.TraceTarget(len(states))
        gl.GeodesicBatch(pipe, i).encode(torch.from_numpy(states).cuda(),
                                         -torch.ones(len(states), dtype=torch.float32,
                                                     device="cuda"), tb)
        torch.cuda.synchronize()
        hp = tb.to_host()
        assert np.all(hb["termination"] > 0)
        for k in hp:
            assert np.array_equal(np.asarray(hb[k]), np.asarray(hp[k])), (i, k)


@pytest.mark.parametrize("refine", [False, True], ids=["f32", "refined"])
def test_shadow_edge_requests_match_the_oracle(gl, oracle, refine):
    """The uneven workload's shadow-edge rays (10% of each scene's pixel
    count as geodesic-batch requests at offsets delta ~ +-log-U[1e-6, 1e-2]
    from the critical curve) of four seed-drawn scenes, traced in one mixed
    whole-tile launch with their images: on 192 seeded requests of each, every
    definite class and event id equals the oracle's (rtol 1e-10, flagged when
    the class changes at 1e-10 / 16) and every settled branch key is exact.
    The f32 trace leaves the requests within its near-critical margin
    Uncertain (a third of them, the offsets below ~1e-5); with the f64 refine
    pass at most 1% stay Uncertain."""
    from gravlume_amd import scene as S
    import torch
    scenes = S.uneven_scenes(count=4, seed=20261206)
    for s in scenes:
        s.width = s.height = 128
        s.refine = refine
    pipe = gl.TracePipeline(scenes)
    items, keep = [], []
    for i, s in enumerate(scenes):
        items.append(("image", i, pipe.create_target(i)))
        states, delta = S.shadow_edge_requests(s, 0.1, seed=700 + i)
        d_states = torch.from_numpy(states).cuda()
        d_dirs = -torch.ones(len(states), dtype=torch.float32, device="cuda")
        bt = gl.TraceTarget(len(states))
        items.append(("batch", i, d_states, d_dirs, bt))
        keep.append((d_states, d_dirs, bt, states, delta))
    pipe.encode_work(items)
    torch.cuda.synchronize()
    rng = np.random.default_rng(17)
    for i, s in enumerate(scenes):
        _, _, bt, states, delta = keep[i]
        h = bt.to_host()
        assert np.all(h["termination"] > 0)
        idx = np.sort(rng.choice(len(states), min(192, len(states)), replace=False))
        st = oracle.spacetime(1.0, s.spin, s.charge, float(s.chart))
        ev = oracle.events(s.escape_radius, s.disk_radii(), horizon_offset=s.horizon_offset)
        dirs = -np.ones(len(idx))
        A = oracle.outcomes_to_arrays(oracle.trace_ordered(
            st, oracle.config_policy(1e-10, 1e-12), ev, states[idx], dirs, threads=16))
        F = oracle.outcomes_to_arrays(oracle.trace_ordered(
            st, oracle.config_policy(1e-10 / 16, 1e-12 / 16), ev, states[idx], dirs, threads=16))
        ok = (A["termination"] == F["termination"]) & (A["candidates"] == F["candidates"])
        t = h["termination"][idx]
        definite = ok & (t != 6)
        print(i, f"a={s.spin:.3f} Q={s.charge:.3f}", "requests", len(states), "flagged",
              int((~ok).sum()), "uncertain", int((ok & (t == 6)).sum()),
              "branch unsettled", int((definite & h["branch_unsettled"][idx]).sum()))
        assert np.mean(t[ok] == 6) <= (0.01 if refine else 0.5)
        assert np.array_equal(t[definite], A["termination"][definite])
        assert np.array_equal(h["candidates"][idx][definite], A["candidates"][definite])
        settled = definite & ~h["branch_unsettled"][idx]
        for key in ("radial_turnings", "equatorial_crossings", "azimuth_winding", "polar_side"):
            assert np.array_equal(h[key][idx][settled], A[key][settled]), key


def test_uneven_workload_launch_matches_the_oracle(gl, oracle):
    """The uneven multi-scene workload itself, as the benchmark builds it: the
    256 scenes of seed 20261206 at their drawn sizes (1.03e8 pixels) and,
    per scene, 10% of its pixel count as shadow-edge requests (seed
    20261207 + scene index, 1.03e7 requests), all in one whole-tile
    work-list launch. Eight seeded pixels and eight seeded requests of every
    scene (2,048 of each) against the oracle at rtol 1e-10: the pooled pixel
    sample with the frame tolerances above; on the requests every definite
    class and event id equals the oracle's and every settled branch key is
    exact, the f32 trace leaving the requests within its near-critical
    margin Uncertain."""
    from gravlume_amd import scene as S
    import torch
    seed = 20261206
    scenes = S.uneven_scenes(count=256, seed=seed)
    pipe = gl.TracePipeline(scenes)
    items, keep = [], []
    for i, s in enumerate(scenes):
        img = pipe.create_target(i)
        items.append(("image", i, img))
        states, _ = S.shadow_edge_requests(s, 0.1, seed=seed + 1 + i)
        d_states = torch.from_numpy(states).cuda()
        d_dirs = -torch.ones(len(states), dtype=torch.float32, device="cuda")
        bt = gl.TraceTarget(len(states))
        items.append(("batch", i, d_states, d_dirs, bt))
        keep.append((img, bt, states, d_states, d_dirs))
    print("rays", sum(s.rays for s in scenes), "requests", sum(len(k[2]) for k in keep), flush=True)
    pipe.encode_work(items)
    torch.cuda.synchronize()
    rng = np.random.default_rng(seed + 4)
    hp, ap, hq, aq = [], [], [], []
    for i, s in enumerate(scenes):
        img, bt, states, _, _ = keep[i]
        h = img.to_host()
        assert np.all(h["termination"] > 0), i
        idx = np.sort(rng.choice(s.rays, 8, replace=False))
        hp.append({k: v[idx] for k, v in h.items()})
        ap.append(_oracle_outcomes(oracle, s, idx, threads=16))
        hb = bt.to_host()
        assert np.all(hb["termination"] > 0), i
        j = np.sort(rng.choice(len(states), 8, replace=False))
        hq.append({k: v[j] for k, v in hb.items()})
        st = oracle.spacetime(1.0, s.spin, s.charge, float(s.chart))
        ev = oracle.events(s.escape_radius, s.disk_radii(), horizon_offset=s.horizon_offset)
        A = oracle.outcomes_to_arrays(oracle.trace_ordered(
            st, oracle.config_policy(1e-10, 1e-12), ev, states[j], -np.ones(8), threads=16))
        F = oracle.outcomes_to_arrays(oracle.trace_ordered(
            st, oracle.config_policy(1e-10 / 16, 1e-12 / 16), ev, states[j], -np.ones(8),
            threads=16))
        A["flagged"] = (A["termination"] != F["termination"]) | (A["candidates"] != F["candidates"])
        aq.append(A)

    def pool(parts):
        return {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}

    H, A = pool(hp), pool(ap)
    n = len(H["termination"])
    print("pixels", _compare(H, A, np.arange(n), dict(TOL, uncertain=1.2e-2)),
          "flagged", int(A["flagged"].sum()))
    H, A = pool(hq), pool(aq)
    ok = ~A["flagged"]
    t = H["termination"]
    definite = ok & (t != 6)
    print("requests: flagged", int((~ok).sum()), "uncertain", int((ok & (t == 6)).sum()),
          "definite", int(definite.sum()))
    assert definite.sum() >= 1024
    assert np.array_equal(t[definite], A["termination"][definite])
    assert np.array_equal(H["candidates"][definite], A["candidates"][definite])
    settled = definite & ~H["branch_unsettled"]
    for key in ("radial_turnings", "equatorial_crossings", "azimuth_winding", "polar_side"):
        assert np.array_equal(H[key][settled], A[key][settled]), key


def test_default_view_produces_only_determinate_rays(gl):
    """The reference's default view at 80x45 (gpu_trace_tests.rs:356-378):
    every ray crosses the horizon or escapes, and both occur."""
    from gravlume_amd import scene as S
    import torch
    sc = S.Scene(spin=0.8, charge=0.0, inclination_deg=60.0, observer_radius=30.0, width=80,
                 height=45, fov_deg=45.0, chart=S.OUTGOING, disk=None, escape_radius=200.0,
                 escape_map=False)
    pipe = gl.TracePipeline([sc])
    tgt = pipe.create_target(0)
    pipe.encode_batch(tgt)
    torch.cuda.synchronize()
    t = tgt.to_host()["termination"]
    assert set(np.unique(t)) == {1, 2}, np.unique(t, return_counts=True)
