"""Autotuner for the persistent trace kernel.

Times the persistent frame on the device for each candidate launch
configuration (register budget of 2, 3 or 4 waves per SIMD, refill interval,
refill threshold; a threshold of 64 refills whole 8x8 tiles only, one tile
per wavefront, which keeps neighbouring pixels together in a wave) with HIP
events around the kernel, and installs the fastest
one on the pipeline. The search is a small exhaustive grid: each candidate
is timed `repeats` times after one warm-up launch and scored by its median.
"""
from __future__ import annotations

import itertools
import json
import statistics

import torch

DEFAULT_GRID = {
    "waves_per_simd": (2, 3, 4),
    "refill_interval": (4, 8, 16, 32),
    "refill_threshold": (8, 16, 32, 64),
}


def autotune(pipeline, targets, scenes=None, grid=None, repeats=3, block_size=256,
             verbose=False) -> dict:
    """Returns {"config": best, "ms": best median ms, "trials": [...]}."""
    grid = dict(DEFAULT_GRID, **(grid or {}))
    pipeline.enable_timing(True)
    trials = []
    best = None
    for waves, interval, thr in itertools.product(grid["waves_per_simd"], grid["refill_interval"],
                                                  grid["refill_threshold"]):
        if thr >= 64 and (waves > 3 or interval != grid["refill_interval"][0]):
            continue  # whole-tile mode: no interval, 2 or 3 waves
        cfg = dict(block_size=block_size, blocks_per_cu=waves, refill_interval=interval,
                   refill_threshold=thr, waves_per_simd=waves)
        pipeline.set_persistent_config(**cfg)
        pipeline.encode_persistent(targets, scenes)
        pipeline.kernel_ms()  # discard the warm-up launch
        times = []
        for _ in range(repeats):
            pipeline.encode_persistent(targets, scenes)
            times.append(pipeline.kernel_ms()[0])
        med = statistics.median(times)
        trials.append({"config": cfg, "ms": med})
        if verbose:
            print(json.dumps(trials[-1]), flush=True)
        if best is None or med < best["ms"]:
            best = {"config": cfg, "ms": med}
    torch.cuda.synchronize()
    pipeline.set_persistent_config(**best["config"])
    return {"config": best["config"], "ms": best["ms"], "trials": trials}
