#!/usr/bin/env python3
"""Throughput of the Kerr-Newman null-geodesic trace on one MI355X.

Workload: one 1920x1080 frame of Kerr a = 0.94 (Q = 0) seen from r = 100 M at
80 deg inclination, FOV 15 deg; adaptive f32 Dormand-Prince 5(4) with
rtol 1e-6 / atol 1e-8; a thin prograde disk [r_ISCO, 30 M] and escape to
r = 1000 M looked up in a seeded 2048x1024 equirectangular RGB map. A step is
one full frame (2,073,600 rays) traced by the persistent kernel; the plain
one-ray-per-lane kernel is timed on the same frame for the ratio.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
                    [--persistent-config WAVES,INTERVAL,THRESHOLD]
                    [--copy-calibration-mib M]

Prints one JSON line. Kernel durations are the mean over the K timed frames,
from HIP events the library records on its launch stream around the main
kernel and the f64 finisher. --persistent-config skips the autotuner (for
profiling runs); --copy-calibration-mib first copies M MiB device to device,
the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-vjp", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-near-critical", action="store_true",
                    help="skip the near-critical batch leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven multi-scene leg")
    ap.add_argument("--no-accel", action="store_true", help="skip the accelerated sky leg")
    ap.add_argument("--no-footprint", action="store_true", help="skip the surface footprint leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--accuracy-rays", type=int, default=32768)
    ap.add_argument("--persistent-config", type=str, default="2,8,64",
                    help="WAVES,INTERVAL,THRESHOLD of the persistent launch (default: whole-tile "
                         "mode, two waves per SIMD, the variant the committed profiles cover)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the persistent launch configuration with the autotuner instead")
    ap.add_argument("--copy-calibration-mib", type=int, default=0)
    ap.add_argument("--only", choices=["near_critical"], default=None,
                    help="run one workload's leg alone and print its JSON (for profiling); the "
                         "default runs the full bench line")
    return ap.parse_args()


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def measured_traffic(path=os.path.join(PROFILES, "k_persistent_tiles_1080p_traffic.md")):
    """HBM bytes per launch of the persistent kernel on this workload, from the
    committed FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel
    calibration recorded there), or None if the summary is absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def kernel_variant(cfg: dict) -> str:
    """Name of the persistent kernel instantiation a configuration launches
    on a one-scene frame."""
    if cfg["refill_threshold"] >= 64:
        return f"k_persistent_tiles<true,{cfg['waves_per_simd']}>"
    return f"k_persistent<true,{cfg['waves_per_simd']}>"


def angle(a, b):
    cr = np.linalg.norm(np.cross(a, b), axis=-1)
    return np.arctan2(cr, np.sum(a * b, axis=-1))


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, gpu_host, seconds, accuracy_rays, seed=20261203):
    """The oracle (CPU f64 DP5 restatement of the reference trace) on the host
    cores: a timed sample at the workload's tolerance, and a tight-tolerance
    sample whose observables measure the f32 error of the GPU frame."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O

    cores = max(1, min(16, os.cpu_count() or 1))
    obs = O.observation(1.0, scene.spin, scene.charge, float(scene.chart),
                        [0.0, *scene.observer_xyz()], [0, 0, 0, 0], [0, 0, 1], 1.0,
                        scene.width, scene.height, np.radians(scene.fov_deg))
    ev = O.events(scene.escape_radius, scene.disk_radii(), horizon_offset=scene.horizon_offset)
    em = O.Emitter()
    rng = np.random.default_rng(seed + 1)
    # timed leg: the workload's own tolerance
    pol = O.config_policy(scene.rtol, scene.atol)
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        idx = rng.integers(0, scene.rays, 1024)
        sxy = O.sight_planes(obs, idx % scene.width, idx // scene.width)
        O.observation_trace_ordered(obs, pol, ev, sxy, em, threads=cores)
        done += len(idx)
    elapsed = time.perf_counter() - t0
    base = {"value": done / elapsed, "unit": "rays/s", "cores": cores, "kind": "port",
            "sample": f"{done} seeded pixels of the 1920x1080 frame, f64 DP5 rtol {scene.rtol:g} "
                      f"atol {scene.atol:g} with the reference policy's 0.5 M step cap (the GPU "
                      f"caps steps at 0.25 r, up to {scene.maximum_step:g} M), {elapsed:.1f} s on "
                      f"{cores} threads",
            "host": {"nproc": os.cpu_count(), "cpu_model": _cpu_model()}}
    # like for like: the same timed sample kind with the GPU trace's step limits
    # (initial step, 256 M maximum, |h| <= 0.25 r, growth cap) instead of the
    # reference policy's 0.5 M cap and growth cap 5
    h0 = scene.initial_step * scene.observer_radius
    gpol = O.gpu_step_policy(scene.rtol, scene.atol, h0, scene.maximum_step,
                             scene.maximum_step_per_radius, scene.step_max_factor)
    done2, t2 = 0, time.perf_counter()
    while time.perf_counter() - t2 < seconds / 2:
        idx = rng.integers(0, scene.rays, 1024)
        sxy = O.sight_planes(obs, idx % scene.width, idx // scene.width)
        O.observation_trace_ordered(obs, gpol, ev, sxy, em, threads=cores)
        done2 += len(idx)
    el2 = time.perf_counter() - t2
    base["gpu_step_policy"] = {
        "rays_per_s": done2 / el2, "cores": cores,
        "sample": f"{done2} seeded pixels, f64 DP5 rtol {scene.rtol:g} atol {scene.atol:g} with the "
                  f"GPU trace's step limits (initial {h0:g} M, at most "
                  f"{scene.maximum_step:g} M and {scene.maximum_step_per_radius:g} r, growth "
                  f"at most {scene.step_max_factor:g}x per step), "
                  f"{el2:.1f} s on {cores} threads"}
    # the small CPU configuration in full: 64x64, a = 0.5, Q = 0.3, rtol 1e-10
    from gravlume_amd import scene as S
    small = S.small_cpu_scene()
    sobs = O.observation(1.0, small.spin, small.charge, float(small.chart),
                         [0.0, *small.observer_xyz()], [0, 0, 0, 0], [0, 0, 1], 1.0, small.width,
                         small.height, np.radians(small.fov_deg))
    sev = O.events(small.escape_radius, small.disk_radii(), horizon_offset=small.horizon_offset)
    allpix = np.arange(small.rays)
    t1 = time.perf_counter()
    O.observation_trace_ordered(sobs, O.config_policy(1e-10, 1e-12), sev,
                                O.sight_planes(sobs, allpix % small.width, allpix // small.width),
                                em, threads=cores)
    base["small_config_full"] = {"rays_per_s": small.rays / (time.perf_counter() - t1),
                                 "sample": "all 4096 pixels of the 64x64 a=0.5 Q=0.3 view, f64 "
                                           "DP5 rtol 1e-10 atol 1e-12"}
    # accuracy leg: f64 reference at rtol 1e-10, and its 1/16 refinement for
    # the near-critical flag (termination changes between the two), on a
    # sample drawn from its own generator (the timed legs above draw as many
    # batches as their time allows)
    idx = np.random.default_rng(seed).choice(scene.rays, accuracy_rays, replace=False)
    sxy = O.sight_planes(obs, idx % scene.width, idx // scene.width)
    ref, _ = O.observation_trace_ordered(obs, O.config_policy(1e-10, 1e-12), ev, sxy, em,
                                         threads=cores)
    fine, _ = O.observation_trace_ordered(obs, O.config_policy(1e-10 / 16, 1e-12 / 16), ev, sxy,
                                          em, threads=cores)
    A = O.outcomes_to_arrays(ref)
    F = O.outcomes_to_arrays(fine)
    flagged = (A["termination"] != F["termination"]) | (A["candidates"] != F["candidates"])
    gt = gpu_host["termination"][idx]
    gc = gpu_host["candidates"][idx]
    ok = ~flagged
    definite = ok & (gt != 6)
    agree = float(np.mean((gt == A["termination"])[ok])) if ok.any() else float("nan")
    agree_ev = float(np.mean(((gt == A["termination"]) & (gc == A["candidates"]))[ok])) if ok.any() else float("nan")
    err = {"rays": int(accuracy_rays), "near_critical_flagged": int(flagged.sum()),
           "termination_agreement": agree, "event_id_agreement": agree_ev,
           "uncertain": int(np.sum(ok & (gt == 6))),
           "definite_class_mismatches": int(np.sum(definite & (gt != A["termination"]))),
           "definite_event_id_mismatches": int(np.sum(definite & (gc != A["candidates"]))),
           "event_residual_max": float(np.max(np.abs(gpu_host["event_residual"][idx][
               definite & np.isin(gt, (1, 2, 3, 7))])))}
    esc = ok & (gt == 2) & (A["termination"] == 2)
    if esc.any():
        a = angle(gpu_host["observable"][idx][esc], A["escape_direction"][esc])
        tt = np.abs(gpu_host["travel_time"][idx][esc] - A["travel_time"][esc])
        err.update(escape_angle_p50=float(np.median(a)), escape_angle_p99=float(np.quantile(a, 0.99)),
                   escape_angle_max=float(a.max()), travel_time_abs_p99=float(np.quantile(tt, 0.99)),
                   travel_time_rel_max=float(np.max(tt / np.abs(A["travel_time"][esc]))))
    srf = ok & (gt == 7) & (A["termination"] == 7)
    if srf.any():
        g = gpu_host["observable"][idx][srf]
        rel = np.abs(g[:, 2] / A["frequency_ratio"][srf] - 1.0)
        err.update(g_rel_p99=float(np.quantile(rel, 0.99)), g_rel_max=float(rel.max()),
                   source_r_abs_max=float(np.max(np.abs(g[:, 0] - A["source_radius"][srf]))))
    err["null_drift_p99"] = float(np.quantile(gpu_host["drift"][idx][:, 0], 0.99))
    # branch key on the definite rays whose branch the trace settles
    same = definite & ~gpu_host["branch_unsettled"][idx]
    err["branch_settled_rays"] = int(same.sum())
    if same.any():
        err["branch_key_mismatches"] = {
            k: int(np.sum((gpu_host[k][idx] != A[k])[same]))
            for k in ("radial_turnings", "equatorial_crossings", "azimuth_winding", "polar_side")}
    return base, err


def vjp_leg(g, S, torch, repeats=5):
    """Adjoint workload: 512x512, a = 0.9, Q = 0.2, inclination 70 deg; loss
    sum (g - g*)^2 over surface pixels with g* traced at (0.92, 0.18, 71 deg);
    checkpoints every 32 accepted steps. Kernel times from the library's
    events: the plain forward frame (trace + f64 finisher) against the
    adjoint's recording forward and its backward."""
    sc = S.adjoint_scene()
    pipe = g.TracePipeline([sc])
    pipe.enable_timing(True)
    tgt = pipe.create_target(0)
    pipe.encode_batch(tgt)
    torch.cuda.synchronize()
    pipe.kernel_ms()
    for _ in range(repeats):
        pipe.encode_batch(tgt)
    torch.cuda.synchronize()
    main_ms, fin_ms, _ = pipe.kernel_ms()
    tsc = S.adjoint_scene(spin=0.92, charge=0.18, inclination_deg=71.0)
    tpipe = g.TracePipeline([tsc])
    ttg = tpipe.create_target(0)
    tpipe.encode_batch(ttg)
    torch.cuda.synchronize()
    n = ttg.rays
    gstar = torch.where((ttg.status & 15) == 7, ttg.observable[2 * n:], torch.zeros_like(ttg.affine))
    fwd, bwd, loss, grads = [], [], 0.0, None
    for _ in range(repeats + 1):
        v = g.VjpTrace(pipe, 0, checkpoint_interval=32)
        torch.cuda.synchronize()
        f_ms = pipe.kernel_ms()[0]
        gg = v.g
        mask = (gg > 0) & (gstar > 0)
        resid = torch.where(mask, gg - gstar, torch.zeros_like(gg))
        loss = float((resid * resid).sum())
        grads = v.backward(2.0 * resid)
        b_ms = pipe.kernel_ms()[0]
        v.close()
        fwd.append(f_ms)
        bwd.append(b_ms)
    fwd, bwd = float(np.median(fwd[1:])), float(np.median(bwd[1:]))
    plain = main_ms + fin_ms
    return {"workload": "512x512 Kerr-Newman a=0.9 Q=0.2 incl 70deg r_obs 100M FOV 20deg, disk "
                        "[r_ISCO, 25M], f32 rtol 1e-6, K=32; loss sum (g-g*)^2, g* at "
                        "(0.92, 0.18, 71deg)",
            "plain_forward_ms": plain, "vjp_forward_ms": fwd, "vjp_backward_ms": bwd,
            "cost_ratio": (fwd + bwd) / plain, "backward_over_forward": bwd / fwd,
            "loss": loss, "grad_a_q_incl": [float(x) for x in grads]}


def near_critical_leg(g, S, torch, n=1_000_000, sample=1024, seed=20261208):
    """Near-critical batch: a = 0.998, Q = 0.05, observer r = 200 M at 85 deg;
    n requests at offsets delta ~ +-log-U[1e-8, 1e-3] from the critical curve
    (f64 initial data), f32 DP5 rtol 1e-6 with the f64 steps inside
    1.05 r_+, the interval capture certificate enabled. Three runs of the
    same batch: the plain batch kernel and the persistent lane-refill queue
    (rays the f32 trace cannot settle leave Uncertain), and the persistent
    queue with the f64 refine pass at rtol 1e-12 (the oracle's tolerance);
    then the whole batch in f64 at rtol 1e-12 (gl_geodesic_batch_trace_f64),
    beside the oracle's own rate at that tolerance on the sample.
    Against the f64 oracle at rtol 1e-12 on a seeded sample, by |delta|
    band: the Uncertain share, the class agreement of definite rays, and the
    winding and travel time of definite rays whose branch is settled."""
    import math
    from gravlume_amd import critical as Cr
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    states, delta, _, _ = Cr.near_critical_states(0.998, 0.05, 200.0, math.radians(85.0), n,
                                                  seed=seed, chart=-1)
    d_states = torch.from_numpy(states).cuda()
    d_dirs = -torch.ones(n, dtype=torch.float32, device="cuda")
    res = {"workload": f"{n} rays, Kerr-Newman a=0.998 Q=0.05, observer r=200M incl 85deg, "
                       "signed offsets delta ~ log-U[1e-8,1e-3] from the critical curve, "
                       "f32 DP5 rtol 1e-6 + f64 steps inside 1.05 r+, capture certificate on; "
                       "refined: near-critical rays traced again in f64 at rtol 1e-12"}
    hosts = {}
    for name, persistent, refine in (("plain", False, False), ("persistent", True, False),
                                     ("persistent_refined", True, True)):
        sc = S.near_critical_scene(capture_certificate=True, refine=refine)
        pipe = g.TracePipeline([sc])
        pipe.enable_timing(True)
        batch = g.GeodesicBatch(pipe, 0)
        t = g.TraceTarget(n)
        batch.encode(d_states, d_dirs, t, persistent=persistent)  # warm-up
        torch.cuda.synchronize()
        pipe.kernel_ms()
        t0 = time.perf_counter()
        batch.encode(d_states, d_dirs, t, persistent=persistent)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        main_ms, f64_ms, _ = pipe.kernel_ms()
        h = t.to_host()
        cert = (h["flags"] & 64) != 0
        res[name] = {"rays_per_s": n / wall, "ms": 1e3 * wall, "kernel_ms": main_ms,
                     "f64_passes_ms": f64_ms, "uncertain": int(np.sum(h["termination"] == 6)),
                     "near_critical": int(h["near_critical"].sum()),
                     "branch_unsettled": int(h["branch_unsettled"].sum()),
                     "refined": int(h["refined"].sum())}
        res[name]["certificate"] = {"accepted": int(cert.sum()), "fallback": int(n - cert.sum()),
                                    "note": "the certificate's proven envelope is |a| <= 0.9; at "
                                            "a = 0.998 every ray falls back to the full trace"}
        hosts[name] = h
        del pipe, t
    # the whole batch in f64 at the oracle's tolerance (gl_geodesic_batch_trace_f64)
    sc64 = S.near_critical_scene(capture_certificate=True)
    pipe = g.TracePipeline([sc64])
    pipe.enable_timing(True)
    batch = g.GeodesicBatch(pipe, 0)
    t = g.TraceTarget(n)
    batch.encode(d_states, d_dirs, t, f64_tolerance=(1e-12, 1e-14))  # warm-up
    torch.cuda.synchronize()
    pipe.kernel_ms()
    t0 = time.perf_counter()
    batch.encode(d_states, d_dirs, t, f64_tolerance=(1e-12, 1e-14))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    main_ms, _, _ = pipe.kernel_ms()
    hosts["f64_batch"] = t.to_host()
    res["f64_batch"] = {"rays_per_s": n / wall, "ms": 1e3 * wall, "kernel_ms": main_ms,
                        "uncertain": int(np.sum(hosts["f64_batch"]["termination"] == 6)),
                        "note": "every request in f64 DP5 at rtol 1e-12 atol 1e-14 from its exact "
                                "initial state"}
    del pipe, t
    res["persistent_over_plain"] = res["persistent"]["rays_per_s"] / res["plain"]["rays_per_s"]
    h = hosts["persistent"]
    res["steps"] = {"accepted_mean": float(h["steps"].mean()), "accepted_max": int(h["steps"].max()),
                    "rejected_mean": float(h["rejected"].mean())}
    hr = hosts["persistent_refined"]
    res["steps_refined"] = {"accepted_mean": float(hr["steps"].mean()),
                            "accepted_max": int(hr["steps"].max())}
    from oracle import oracle as O
    sc = S.near_critical_scene()
    idx = np.random.default_rng(seed + 1).choice(n, sample, replace=False)
    cores = max(1, min(16, os.cpu_count() or 1))
    st = O.spacetime(1.0, sc.spin, sc.charge, float(sc.chart))
    ev = O.events(sc.escape_radius, None, horizon_offset=sc.horizon_offset)
    dirs = -np.ones(sample)
    t0 = time.perf_counter()
    ref = O.outcomes_to_arrays(O.trace_ordered(st, O.config_policy(1e-12, 1e-14), ev, states[idx],
                                               dirs, threads=cores))
    t_ref = time.perf_counter() - t0
    fine = O.outcomes_to_arrays(O.trace_ordered(st, O.config_policy(1e-12 / 16, 1e-14 / 16), ev,
                                                states[idx], dirs, threads=cores))
    t0 = time.perf_counter()
    O.trace_ordered(st, O.config_policy(sc.rtol, sc.atol), ev, states[idx], dirs, threads=cores)
    t_work = time.perf_counter() - t0
    res["cpu_baseline_sample"] = {"rays_per_s": sample / t_work, "cores": cores, "kind": "port",
                                  "sample": f"{sample} seeded requests of the batch, f64 DP5 rtol "
                                            f"{sc.rtol:g} atol {sc.atol:g}, 0.5 M step cap",
                                  "rays_per_s_at_rtol_1e-12": sample / t_ref}
    ok = ref["termination"] == fine["termination"]
    ad = np.abs(delta[idx])
    cmp = {}
    for name in ("persistent", "persistent_refined", "f64_batch"):
        hh = hosts[name]
        t = hh["termination"][idx]
        bands = {}
        for lo, hi in ((1e-8, 1e-6), (1e-6, 1e-5), (1e-5, 1e-3)):
            m = ok & (ad >= lo) & (ad < hi)
            d = m & (t != 6)
            s_ = d & ~hh["branch_unsettled"][idx]
            tt = np.abs(hh["travel_time"][idx][s_] / ref["travel_time"][s_] - 1.0)
            bands[f"{lo:g}-{hi:g}"] = {
                "rays": int(m.sum()), "uncertain_share": float(np.mean(t[m] == 6)) if m.any() else None,
                "definite": int(d.sum()),
                "class_agreement_definite": float(np.mean(t[d] == ref["termination"][d])) if d.any() else None,
                "branch_settled": int(s_.sum()),
                "winding_agreement_settled": float(np.mean(hh["azimuth_winding"][idx][s_] ==
                                                           ref["azimuth_winding"][s_])) if s_.any() else None,
                "travel_time_rel_max_settled": float(tt.max()) if s_.any() else None}
        cmp[name] = bands
    res["vs_f64_rtol_1e-12"] = {"sample": sample, "oracle_flagged": int((~ok).sum()),
                                "by_abs_delta": cmp}
    return res


def uneven_leg(g, S, torch, count=256, seed=20261206, jitter=0.1):
    """Uneven multi-scene batch in one launch: `count` seed-drawn scenes
    (a ~ U[0, 0.998], Q ~ U[0, 0.9 sqrt(1 - a^2)], inclination U[5, 89] deg,
    r_obs U[30, 200] M, side in {128, 256, 512, 768, 1024}, FOV U[8, 30] deg,
    disk [r_ISCO, 20 M], escape 1000 M), every pixel traced, plus per scene
    `jitter` x its pixel count of extra rays seeded next to its shadow edge
    (offsets +-log-U[1e-6, 1e-2] from the critical curve) as geodesic-batch
    requests. One persistent launch over the whole work list (lane refill and
    whole tiles) against one plain launch per image and per batch."""
    scenes = S.uneven_scenes(count=count, seed=seed)
    pipe = g.TracePipeline(scenes)
    pipe.enable_timing(True)
    items, keep = [], []
    image_rays, batch_rays = 0, 0
    per_scene = []
    for i, sc in enumerate(scenes):
        img = pipe.create_target(i)
        items.append(("image", i, img))
        states, _ = S.shadow_edge_requests(sc, jitter, seed=seed + 1 + i)
        d_states = torch.from_numpy(states).cuda()
        d_dirs = -torch.ones(len(states), dtype=torch.float32, device="cuda")
        bt = g.TraceTarget(len(states))
        items.append(("batch", i, d_states, d_dirs, bt))
        keep.append((d_states, d_dirs, bt, img))
        image_rays += sc.rays
        batch_rays += len(states)
        per_scene.append([sc.width, len(states)])
    rays = image_rays + batch_rays
    res = {"workload": f"{count} seed-drawn scenes (seed {seed}), {image_rays} pixel rays + "
                       f"{batch_rays} shadow-edge rays, f32 DP5 rtol 1e-6, one launch",
           "rays": rays}
    for name, cfg in (("persistent_lane_refill", dict(refill_threshold=16, refill_interval=8)),
                      ("persistent_whole_tile", dict(refill_threshold=64, refill_interval=8))):
        pipe.set_persistent_config(block_size=256, blocks_per_cu=2, waves_per_simd=2, **cfg)
        pipe.encode_work(items)  # warm-up
        torch.cuda.synchronize()
        pipe.kernel_ms()
        t0 = time.perf_counter()
        pipe.encode_work(items)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        main_ms, fin_ms, _ = pipe.kernel_ms()
        st = pipe.queue_stats()
        res[name] = {"rays_per_s": rays / wall, "ms": 1e3 * wall, "kernel_ms": main_ms,
                     "f64_finisher_ms": fin_ms,
                     "lane_utilization": st["attempted_steps"] / max(1, st["lane_slots"])}
    # one plain launch per image and per batch, back to back on one stream
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, (d_states, d_dirs, bt, img) in enumerate(keep):
        pipe.encode_batch(img)
        g.GeodesicBatch(pipe, i).encode(d_states, d_dirs, bt)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    res["plain_separate_launches"] = {"rays_per_s": rays / wall, "ms": 1e3 * wall}
    best = max(res["persistent_lane_refill"]["rays_per_s"], res["persistent_whole_tile"]["rays_per_s"])
    res["persistent_over_plain"] = best / res["plain_separate_launches"]["rays_per_s"]
    res["per_scene_side_and_edge_rays"] = per_scene
    return res


def footprint_leg(g, S, torch, repeats=3):
    """Source-chart footprint of every pixel of the reference's default view
    (a = 0.8, ingoing chart, r_obs = 30 M at 60 deg, 1280x720, FOV 45 deg,
    surface [6, 20] M): five real traces per pixel (+-0.25-pixel central
    differences, gl_trace_surface_footprint) against one adjoint forward
    trace plus two backward passes (gl_vjp_surface_footprint), and their
    agreement on the pixels both resolve."""
    import ctypes as C
    sc = S.Scene(spin=0.8, charge=0.0, inclination_deg=60.0, observer_radius=30.0, width=1280,
                 height=720, fov_deg=45.0, chart=S.INGOING, disk=(6.0, 20.0), escape_radius=200.0,
                 escape_map=False, emission="bolometric")
    pipe = g.TracePipeline([sc])
    n = sc.rays
    lib = pipe._lib
    jac = [torch.zeros(4 * n, dtype=torch.float32, device="cuda") for _ in range(2)]
    meta = [torch.zeros(n, dtype=torch.int32, device="cuda") for _ in range(2)]

    def five_ray():
        g._capi.check(lib.gl_trace_surface_footprint(pipe.handle, 0, None, n, jac[0].data_ptr(),
                                                     meta[0].data_ptr(), None))

    def adjoint():
        vt = g.VjpTrace(pipe, 0)
        out = vt.target.c_struct()
        g._capi.check(lib.gl_vjp_surface_footprint(vt._handle, C.byref(out), C.c_void_p(jac[1].data_ptr()),
                                                   C.c_void_p(meta[1].data_ptr()), None))
        torch.cuda.synchronize()
        vt.close()

    res = {"workload": "surface footprint of every pixel, default view 1280x720, a=0.8 Q=0 ingoing "
                       "chart, observer r=30M incl 60deg, FOV 45deg, surface [6, 20]M, f32 rtol 1e-6"}
    for name, fn in (("five_ray", five_ray), ("adjoint", adjoint)):
        fn()
        torch.cuda.synchronize()
        t = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
        res[name] = {"ms": 1e3 * float(np.median(t)), "pixels_per_s": n / float(np.median(t))}
    m = [x.cpu().numpy().view(np.uint32) for x in meta]
    J = [x.cpu().numpy().astype(np.float64).reshape(n, 2, 2) for x in jac]
    both = ((m[0] & 1) != 0) & ((m[1] & 1) != 0)
    err = np.max(np.abs(J[0][both] - J[1][both]), axis=(1, 2)) / np.max(np.abs(J[1][both]), axis=(1, 2))
    res["resolved"] = {"five_ray": int(np.sum(m[0] & 1)), "adjoint": int(np.sum(m[1] & 1)),
                       "both": int(both.sum())}
    res["max_norm_rel_diff"] = {"p50": float(np.median(err)), "p99": float(np.quantile(err, 0.99)),
                                "max": float(err.max())}
    res["parity_agreement"] = float(np.mean(((m[0] >> 1) & 3)[both] == ((m[1] >> 1) & 3)[both]))
    res["adjoint_over_five_ray_speed"] = res["five_ray"]["ms"] / res["adjoint"]["ms"]
    return res


def accelerated_sky_leg(g, S, torch, repeats=5):
    """Image-space accelerators on a sky-only view: the reference's default
    transfer profile (a = 0.8, outgoing chart, r_obs = 30 M at 60 deg, FOV
    45 deg) at 1920x1080 with the capture certificate and the escape map.
    The full trace (plain and persistent) against the accelerated sky trace
    (escape-direction stencil), with the exactness of its classes and the
    chord error of its reconstructed directions; then the shadow-edge
    four-sample refinement of the traced frame."""
    sc = S.Scene(spin=0.8, charge=0.0, inclination_deg=60.0, observer_radius=30.0, width=1920,
                 height=1080, fov_deg=45.0, chart=S.OUTGOING, disk=None, escape_radius=200.0,
                 capture_certificate=True)
    pipe = g.TracePipeline([sc])
    pipe.enable_timing(True)
    full, full_p, acc = pipe.create_target(0), pipe.create_target(0), pipe.create_target(0)

    def run(fn):
        fn()  # warm-up
        torch.cuda.synchronize()
        pipe.kernel_ms()
        t0 = time.perf_counter()
        for _ in range(repeats):
            fn()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / repeats
        main_ms, fin_ms, n = pipe.kernel_ms()
        return {"rays_per_s": sc.rays / wall, "ms": 1e3 * wall, "kernel_ms": main_ms,
                "f64_finisher_ms": fin_ms}

    res = {"workload": "sky-only view 1920x1080, a=0.8 Q=0 outgoing chart, observer r=30M incl "
                       "60deg, FOV 45deg, escape r=200M into the map, capture certificate, f32 DP5 "
                       "rtol 1e-6"}
    res["full_plain"] = run(lambda: pipe.encode_batch(full))
    res["full_persistent"] = run(lambda: pipe.encode_persistent([full_p]))
    res["accelerated"] = run(lambda: pipe.encode_accelerated(acc))
    hf, ha = full.to_host(), acc.to_host()  # against the plain full trace
    rec = (ha["flags"] & 128) != 0
    esc = hf["termination"] == 2
    chord = np.linalg.norm(ha["observable"][esc].astype(np.float64) - hf["observable"][esc], axis=1)
    res["reconstructed_fraction"] = float(rec.mean())
    res["termination_agreement"] = float(np.mean(ha["termination"] == hf["termination"]))
    res["candidate_agreement"] = float(np.mean(ha["candidates"] == hf["candidates"]))
    res["chord_max"] = float(chord.max())
    res["chord_p99_reconstructed"] = float(np.quantile(
        np.linalg.norm(ha["observable"][rec].astype(np.float64) - hf["observable"][rec], axis=1), 0.99)) \
        if rec.any() else 0.0
    best_full = max(res["full_plain"]["rays_per_s"], res["full_persistent"]["rays_per_s"])
    res["accelerated_over_full"] = res["accelerated"]["rays_per_s"] / best_full
    pipe.encode_batch(full)
    torch.cuda.synchronize()
    pipe.kernel_ms()
    t0 = time.perf_counter()
    pipe.refine_shadow(full)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    edges = pipe.shadow_edges()
    main_ms, _, _ = pipe.kernel_ms()
    res["shadow_refine"] = {"edge_pixels": int(edges), "ms": 1e3 * wall, "kernel_ms": main_ms,
                            "subsample_rays_per_s": 4 * edges / (main_ms * 1e-3) if main_ms else None}
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (rays are independent, so "
              "a frame or a work list would shard by whole scenes or tile ranges, one process per "
              "GPU, with no collective inside the trace)", file=sys.stderr)
        sys.exit(2)
    import torch
    import gravlume_amd as g
    from gravlume_amd import scene as S
    from gravlume_amd.autotune import autotune

    if args.only == "near_critical":
        print(json.dumps({"near_critical": near_critical_leg(g, S, torch)}), flush=True)
        return
    scene = S.single_frame_1080p()
    pipe = g.TracePipeline([scene])
    pipe.enable_timing(True)
    rhs_flops, attempt_flops = g.flop_constants()
    target = pipe.create_target(0)
    plain_target = pipe.create_target(0)
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if not args.autotune:
        w, i, t = (int(v) for v in args.persistent_config.split(","))
        pipe.set_persistent_config(block_size=256, blocks_per_cu=w, refill_interval=i,
                                   refill_threshold=t, waves_per_simd=w)
        tuned = {"config": pipe.persistent_config()}
    else:
        tuned = autotune(pipe, [target])
    for _ in range(args.warmup):
        pipe.encode_persistent([target])
        pipe.encode_batch(plain_target)
    torch.cuda.synchronize()

    def timed(fn):
        # wall clock over exactly `steps` frames, bracketed by synchronization;
        # the library's events give the mean kernel durations of those frames
        torch.cuda.synchronize()
        pipe.kernel_ms()  # start a new count
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        main_ms, fin_ms, launches = pipe.kernel_ms()
        assert launches == args.steps, (launches, args.steps)
        return wall, main_ms, fin_ms

    # serial frames: each frame's f64 passes follow its trace on one stream
    t_serial, k_serial, f_serial = timed(lambda: pipe.encode_persistent([target]))
    stats = pipe.queue_stats()
    # frame sequence with the f64 overlap: frame N's f64 passes run on the
    # plan's second stream beside frame N+1's trace, the frames alternating
    # between two targets; every frame is complete when the timed region's
    # closing synchronize returns (outputs bit-identical to serial frames)
    target_b = pipe.create_target(0)
    frames = (target, target_b)
    count = [0]

    def overlapped():
        pipe.encode_persistent([frames[count[0] % 2]])
        count[0] += 1

    pipe.set_f64_overlap(True)
    for _ in range(2):
        overlapped()
    t_pers, k_pers, f_pers = timed(overlapped)
    pipe.set_f64_overlap(False)
    # host-inclusive rate: each frame's 68 B/ray of output planes copied to
    # pinned host memory after the trace (PCIe), frames back to back
    planes = ("status", "steps", "rejected", "observable", "travel_time", "affine", "drift",
              "branch", "radiance", "rgba16f")
    pinned = {k: torch.empty(getattr(target, k).shape, dtype=getattr(target, k).dtype,
                             pin_memory=True) for k in planes}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pipe.encode_persistent([target])
        for k in planes:
            pinned[k].copy_(getattr(target, k), non_blocking=True)
    torch.cuda.synchronize()
    t_pcie = time.perf_counter() - t0
    # the same, pipelined: frame N's planes go to pinned host memory on a copy
    # stream while frame N+1 traces (two targets and two pinned sets
    # alternating; a target is written again only after its copy)
    pinned_b = {k: torch.empty_like(v, pin_memory=True) for k, v in pinned.items()}
    sets = ((target, pinned), (target_b, pinned_b))
    main, copier = torch.cuda.current_stream(), torch.cuda.Stream()
    copied = [None, None]

    def copy_out(k):
        ready = torch.cuda.Event()
        ready.record(main)
        copier.wait_event(ready)
        with torch.cuda.stream(copier):
            for name in planes:
                sets[k][1][name].copy_(getattr(sets[k][0], name), non_blocking=True)
        copied[k] = torch.cuda.Event()
        copied[k].record(copier)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        k = i % 2
        if copied[k] is not None:
            main.wait_event(copied[k])
        pipe.encode_persistent([sets[k][0]])
        copy_out(k)
    torch.cuda.synchronize()
    t_pcie_pipe = time.perf_counter() - t0
    # the same persistent frame captured once as a hipGraph and replayed
    graph = pipe.frame_graph([target])
    graph.launch()
    graph.stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.launch()
    graph.stream.synchronize()
    t_graph = time.perf_counter() - t0
    torch.cuda.synchronize()
    t_plain, k_plain, f_plain = timed(lambda: pipe.encode_batch(plain_target))
    host = target.to_host()
    plain_host = plain_target.to_host()
    rays = scene.rays
    value = rays * args.steps / t_pers
    plain_value = rays * args.steps / t_plain
    attempts = int(host["steps"].sum() + host["rejected"].sum())
    flops = attempts * attempt_flops
    achieved = flops / (k_pers * 1e-3) / 1e12
    peak = 157.3
    agree_plain = float(np.mean(host["termination"] == plain_host["termination"]))

    line = {
        "metric": "rays/s @1080p Kerr a=0.94, plain vs persistent; VJP cost ratio; err vs f64",
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_pers / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "single frame 1920x1080, Kerr a=0.94 Q=0, observer r=100M incl 80deg, "
                               "FOV 15deg, f32 DP5(4) rtol 1e-6 atol 1e-8, disk [r_ISCO, 30M], "
                               "escape r=1000M into a 2048x1024 seeded RGB map",
                   "rays_per_step": rays, "kernel": kernel_variant(tuned["config"])},
        "serial_rays_per_s": rays * args.steps / t_serial,
        "plain_rays_per_s": plain_value,
        "host_inclusive_rays_per_s": rays * args.steps / t_pcie,
        "host_inclusive_pipelined_rays_per_s": rays * args.steps / t_pcie_pipe,
        "graph_rays_per_s": rays * args.steps / t_graph,
        "persistent_over_plain": (rays * args.steps / t_serial) / plain_value,
        "overlap_gain": t_serial / t_pers,
        "kernel_ms": {"persistent": k_pers, "persistent_f64_finisher": f_pers,
                      "persistent_serial": k_serial, "persistent_serial_f64_finisher": f_serial,
                      "plain": k_plain, "plain_f64_finisher": f_plain},
        "plain_persistent_termination_agreement": agree_plain,
        "steps_per_ray": {"accepted_mean": float(host["steps"].mean()),
                          "rejected_mean": float(host["rejected"].mean()),
                          "accepted_max": int(host["steps"].max())},
        "lane_utilization": stats["attempted_steps"] / max(1, stats["lane_slots"]),
        "persistent_config": tuned["config"],
        "terminations": {int(k): int(v) for k, v in zip(*np.unique(host["termination"], return_counts=True))},
        "surface_texels": {
            "radiance": int(np.sum((host["termination"] == 7) & (host["radiance"][:, 0] >= 0))),
            "not_representable": int(np.sum((host["termination"] == 7) & (host["radiance"][:, 0] < 0))),
            "model": f"{scene.emission} T_6 = {scene.six_temperature():.6g} K, I_6 = {scene.intensity_at_six}, vacuum",
        },
        "seeds": {"escape_map": 20261203, "accuracy_sample": 20261203, "cpu_baseline_sample": 20261204,
                  "uneven_scenes": 20261206, "uneven_shadow_edge_rays": "20261207 + scene index",
                  "near_critical_offsets": 20261208, "near_critical_oracle_sample": 20261209,
                  "footprint_and_vjp": "none (every pixel of fixed scenes)"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "frac_of_nonpacked_fp32": achieved / 78.6,
                     "frac_serial_frames": flops / (k_serial * 1e-3) / 1e12 / peak,
                     "traffic": measured_traffic(),
                     "traffic_from": "profiles/k_persistent_tiles_1080p_traffic.md: rocprofv3 --pmc "
                                     "FETCH_SIZE and WRITE_SIZE of this kernel variant on this "
                                     "frame, scaled by a 256 MiB copy (not measured in this run)",
                     "flops_per_launch": flops, "flops_per_attempt": attempt_flops,
                     "flops_per_rhs": rhs_flops, "attempts_per_launch": attempts},
        "cpu_baseline": None,
    }
    if not args.no_vjp:
        line["vjp"] = vjp_leg(g, S, torch)
    if not args.no_near_critical:
        line["near_critical"] = near_critical_leg(g, S, torch)
    if not args.no_uneven:
        line["uneven_scenes"] = uneven_leg(g, S, torch)
    if not args.no_accel:
        line["accelerated_sky"] = accelerated_sky_leg(g, S, torch)
    if not args.no_footprint:
        line["surface_footprint"] = footprint_leg(g, S, torch)
    if not args.no_cpu:
        base, err = cpu_baseline(scene, host, args.cpu_seconds, args.accuracy_rays)
        line["cpu_baseline"] = base
        line["error_vs_f64"] = err
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
