"""Autotuner of the batched scene build.

Picks, by timing candidates on the scene itself, the settings that do not
change the plain result (grid sizes of the level kernels, in blocks per
compute unit, and the largest node whose subtree one wavefront builds) and
the per-node-size dispatch of the learned method (the
smallest node that takes the learned path, the window half-width), keeping
only learned candidates whose scene SAH stays within `sah_tolerance` of the
plain trees'.  Timings are graph replays (`GpuScene.record_graph`), medians of
`reps` replays.  Results are cached per device and scene signature in a JSON
file ($MGBVH_TUNE_CACHE, default ~/.cache/mgbvh/autotune.json).
"""
from __future__ import annotations

import hashlib
import json
import os
import statistics

import numpy as np

from . import _native as N

GRID_CANDIDATES = [
    dict(bin_blocks_per_cu=128, partition_blocks_per_cu=192, sweep_blocks_per_cu=4,
         small_blocks_per_cu=64, partition_waves=4, bin_waves=4),
    dict(bin_blocks_per_cu=64, partition_blocks_per_cu=128, sweep_blocks_per_cu=4,
         small_blocks_per_cu=32, partition_waves=4, bin_waves=4),
    dict(bin_blocks_per_cu=128, partition_blocks_per_cu=256, sweep_blocks_per_cu=4,
         small_blocks_per_cu=128, partition_waves=4, bin_waves=4),
    dict(bin_blocks_per_cu=32, partition_blocks_per_cu=64, sweep_blocks_per_cu=4,
         small_blocks_per_cu=16, partition_waves=4, bin_waves=4),
]
# nodes of 33..this many objects built one wavefront per subtree (1 = off)
MEDIUM_CANDIDATES = [512, 1024, 2048]
LEARNED_CANDIDATES = [
    dict(learned_min_objects=1024, learned_window=3, learned_guard=0.02),
    dict(learned_min_objects=2048, learned_window=2, learned_guard=0.02),
    dict(learned_min_objects=8192, learned_window=2, learned_guard=0.02),
    dict(learned_min_objects=32768, learned_window=2, learned_guard=0.02),
    dict(learned_min_objects=131072, learned_window=2, learned_guard=0.02),
]


def cache_path() -> str:
    return os.environ.get("MGBVH_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "mgbvh", "autotune.json"))


def scene_signature(scene) -> str:
    h = hashlib.sha1()
    h.update(np.asarray(scene.counts, np.int64).tobytes())
    h.update(np.int64(scene.n_instances).tobytes())
    return h.hexdigest()[:16]


def _time_graph(scene, reps: int) -> float:
    scene.build()  # first replay warms caches and allocations
    ms = []
    for _ in range(reps):
        scene.build()
        ms.append(scene.last_graph_ms)
    return statistics.median(ms)


def autotune(ctx: N.Context, meshes, transforms, instance_mesh, reps: int = 3,
             sah_tolerance: float = 0.01, use_cache: bool = True, log=None) -> dict:
    """Returns {"tuning": {...}, "learned": {...}, "ms": ..., "plain_sah": ...,
    "learned_sah": ..., "trials": [...]} and applies the tuning to `ctx`."""
    from .scene import GpuScene

    probe = GpuScene(ctx, meshes, transforms, instance_mesh, mode="batched", build=False)
    key = f"{ctx.get_tuning()['compute_units']}cu/{scene_signature(probe)}/" \
          f"{'model' if ctx_has_model(ctx) else 'nomodel'}"
    path = cache_path()
    if use_cache and os.path.exists(path):
        try:
            cached = json.load(open(path)).get(key)
        except (OSError, ValueError):
            cached = None
        if cached:
            ctx.set_tuning(**cached["tuning"])
            probe.free()
            return dict(cached, cached=True)
    trials = []
    # 1. grid sizes, plain split rule (the result does not depend on them)
    best_grid, best_ms = None, float("inf")
    for cand in GRID_CANDIDATES:
        cand = dict(cand, medium_max_objects=1)
        ctx.set_tuning(**cand)
        probe.record_graph()  # grid sizes are baked into the graph
        ms = _time_graph(probe, reps)
        trials.append({"tuning": cand, "ms": ms})
        if log:
            log(f"autotune grid {cand}: {ms:.2f} ms")
        if ms < best_ms:
            best_grid, best_ms = cand, ms
    # 2. the largest node built one wavefront per subtree (results unchanged)
    for med in MEDIUM_CANDIDATES:
        cand = dict(best_grid, medium_max_objects=med)
        ctx.set_tuning(**cand)
        probe.record_graph()
        ms = _time_graph(probe, reps)
        trials.append({"tuning": cand, "ms": ms})
        if log:
            log(f"autotune medium_max_objects {med}: {ms:.2f} ms")
        if ms < best_ms:
            best_grid, best_ms = cand, ms
    ctx.set_tuning(**best_grid)
    probe.record_graph()
    probe.build()
    plain_sah = probe.scene_sah()
    probe.free()
    result = {"tuning": best_grid, "learned": None, "ms": best_ms, "plain_sah": plain_sah,
              "learned_sah": None}
    # 3. learned dispatch, if a model is attached
    if ctx_has_model(ctx):
        best_l, best_lms = None, float("inf")
        for cand in LEARNED_CANDIDATES:
            sc = GpuScene(ctx, meshes, transforms, instance_mesh, mode="batched", build=False,
                          method=N.METHOD_LEARNED, learned=cand)
            sc.record_graph()
            ms = _time_graph(sc, reps)
            sah = sc.scene_sah()
            sc.free()
            ok = sah <= plain_sah * (1.0 + sah_tolerance)
            trials.append({"learned": cand, "ms": ms, "sah_ratio": sah / plain_sah, "ok": ok})
            if log:
                log(f"autotune learned {cand}: {ms:.2f} ms, SAH ratio {sah / plain_sah:.4f}")
            if ok and ms < best_lms:
                best_l, best_lms, best_sah = cand, ms, sah
        if best_l is not None:
            result.update(learned=best_l, ms=best_lms, learned_sah=best_sah)
    result["trials"] = trials
    if use_cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            data = json.load(open(path)) if os.path.exists(path) else {}
            data[key] = {k: v for k, v in result.items() if k != "trials"}
            json.dump(data, open(path, "w"), indent=1)
        except (OSError, ValueError):
            pass
    return dict(result, cached=False)


def ctx_has_model(ctx: N.Context) -> bool:
    return getattr(ctx, "model", None) is not None
