#!/usr/bin/env python3
"""BVH build throughput on one MI355X.

Workloads:
* ``scene`` (default): the 2,000-mesh scene (log-uniform 100..200,000
  triangles per mesh, half soups and half heightfields, ~52M triangles) with
  10,000 instances.  A step is one whole scene build: every BLAS, the instance
  world boxes and the TLAS, device-resident input -> device-resident nodes.
  `value` is the batched builder; the per-mesh plain builder (one build call
  per mesh, as the reference builds one BVH per mesh) is timed beside it.
* ``heightfield``: one 1024x512-cell heightfield mesh (1,048,576 triangles),
  a step is one plain build through mgbvh_build_tris_device.
* ``soup``: one 65,536-triangle soup.

Prints one JSON line with Mtri/s, the dominant kernel's HBM roofline fraction
(timed live with HIP events on the build stream) and a CPU baseline (the
oracle, a C restatement of marigold's builder, timed on a bounded sample on
this host).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
REC_BYTES = 28.0  # the builder's object record: box (6 floats) + object index


def tree_terms(nodes: np.ndarray, lens: np.ndarray):
    """Sums over the built tree used by the algorithmic byte counts:
    s_large = object counts summed over internal nodes with > 32 objects (each
    such level reads and writes every record once), n_large = those nodes,
    s_small = objects in the roots of <= 32-object subtrees, n_small_nodes =
    nodes written by the small-subtree kernel."""
    child = nodes["child_node"].astype(np.int64)
    nn = len(nodes)
    parent_len = np.full(nn, np.iinfo(np.int64).max, dtype=np.int64)
    internal = np.nonzero(child)[0]
    parent_len[child[internal]] = lens[internal]
    parent_len[child[internal] + 1] = lens[internal]
    large = (lens > 32) & (child != 0)
    small_root = (lens <= 32) & (parent_len > 32)
    return {
        "s_large": int(lens[large].astype(np.int64).sum()),
        "n_large": int(large.sum()),
        "s_small": int(lens[small_root].astype(np.int64).sum()),
        "n_small_nodes": int(nn - large.sum()),
    }


def kernel_bytes(kind: str, n: int, nverts: int, vstride: int, t: dict) -> float:
    """Algorithmic HBM bytes per build of one kernel class (each byte the
    algorithm must move once; scratch re-reads by the implementation excluded)."""
    if kind == "setup":
        return 12.0 * n + 12.0 * nverts + REC_BYTES * n      # indices, positions, object records
    if kind == "bin":
        return REC_BYTES * t["s_large"]                       # one record read per object per level
    if kind == "partition":
        return 2 * REC_BYTES * t["s_large"]                   # record read + write
    if kind == "sweep":
        return t["n_large"] * (3 * 32 * 7 * 4 + 32.0)        # histogram read + node write
    if kind == "small":  # records, node + len writes, perm write, permuted triangle read + write
        return REC_BYTES * n + 36.0 * t["n_small_nodes"] + 4.0 * n + 24.0 * n
    return 0.0


def survey_bytes(n: int, nverts: int, s_large: int, n_small_objects: int) -> float:
    """Algorithmic bytes of a batched BLAS build by the path's byte model:
    per BLAS 12N (indices) + 12V (positions) + 24N (object
    box write) + 32 per object of every internal node with > 32 objects +
    28 per object of every <= 32-object subtree root + 32 (2N - 1) (node
    write) + 24N (permuted triangle read + write), summed over the meshes."""
    return (12.0 * n + 12.0 * nverts + 24.0 * n + 32.0 * s_large + 28.0 * n_small_objects
            + 32.0 * (2 * n) + 24.0 * n)


def measured_valu(kernel: str, path: str = os.path.join(ROOT, "profiles",
                                                        "small_subtree_sq_counters.md")):
    """VALU issue of `kernel` from the committed SQ counter summary, or None:
    {"insts_per_launch", "issue_frac"} (wave64 VALU instructions x 2 SIMD cycles
    each, over the SIMD cycles of the launch)."""
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 4 and f[0] == "valu_per_launch" and f[1] == kernel:
                return {"insts_per_launch": float(f[2]), "issue_frac": float(f[3])}
    except OSError:
        pass
    return None


def measured_traffic(kernel: str, path: str = os.path.join(ROOT, "profiles",
                                                           "scene_build_traffic.md")):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), or None."""
    try:
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
                return int(f[2])
    except OSError:
        pass
    return None


def terms_from_stats(st: dict) -> dict:
    """The same sums from the builder's own work counters."""
    return {"s_large": int(st["large_object_visits"]), "n_large": int(st["large_nodes"]),
            "n_small_nodes": int(st["node_count"]) - int(st["large_nodes"])}


def workload(name: str):
    from marigold_amd import scenes
    if name == "heightfield":
        pos, tris = scenes.heightfield(1024, 512, seed=2)
        desc = "plain GPU build, 1024x512-cell heightfield (1,048,576 tris), 6-octave value noise"
    elif name == "soup":
        pos, tris = scenes.soup(65536, seed=1)
        desc = "plain GPU build, 65,536-tri random soup"
    else:
        raise SystemExit(f"unknown workload {name}")
    return pos, tris, desc


def cpu_baseline(pos, tris, budget_s: float = 12.0):
    from oracle import oracle as O  # CPU baseline only
    ab = O.tri_aabbs(pos, tris)
    times = []
    t_end = time.time() + budget_s
    while True:
        t0 = time.perf_counter()
        O.build(ab, max_depth=32)
        times.append(time.perf_counter() - t0)
        if time.time() > t_end or len(times) >= 10:
            break
    med = statistics.median(times)
    return {"value": tris.shape[0] / med / 1e6, "unit": "Mtri/s", "cores": 1, "kind": "port",
            "sample": f"{len(times)} full single-threaded builds of the same mesh by the C restatement "
                      f"of marigold bvh (median {med:.3f} s; AABB setup excluded)"}


def host_cores() -> dict:
    """What this host shows: logical CPUs, the ones this process may run on,
    and the CPU share given to one GPU's jobs (OMP_NUM_THREADS)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    share = os.environ.get("OMP_NUM_THREADS")
    return {"cpu_count": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": int(share) if share and share.isdigit() else None}


def cpu_pool_threads() -> int:
    """Threads of the per-mesh pool: the CPU share of this GPU's jobs
    (OMP_NUM_THREADS when set), else the cores this process may use."""
    hc = host_cores()
    return hc["omp_num_threads"] or hc["affinity"] or hc["cpu_count"] or 1


def cpu_baseline_scene(meshes, budget_s: float = 15.0, threads: int = 0):
    """Per-mesh oracle builds on a thread pool (one mesh per task, as rayon's
    par_iter over meshes in model.rs:76-95, with the rayon pool's one worker
    per core), over a random sample of meshes until the budget is spent."""
    from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
    from oracle import oracle as O  # CPU baseline only
    threads = threads or cpu_pool_threads()
    rng = np.random.default_rng(0)
    order = rng.permutation(len(meshes))
    done_tris = 0
    t0 = time.perf_counter()
    deadline = t0 + budget_s

    def one(i):
        p, t = meshes[i]
        O.build(O.tri_aabbs(p, t), max_depth=32)
        return t.shape[0]

    n_done = 0
    with ThreadPoolExecutor(threads) as ex:
        it = iter(order)
        pending = set()
        for _ in range(threads):
            i = next(it, None)
            if i is not None:
                pending.add(ex.submit(one, i))
        while pending:
            # every finished task frees its thread for the next mesh at once
            done, pending = wait(pending, return_when=FIRST_COMPLETED)
            for f in done:
                done_tris += f.result()
                n_done += 1
                if time.perf_counter() < deadline:
                    i = next(it, None)
                    if i is not None:
                        pending.add(ex.submit(one, i))
    wall = time.perf_counter() - t0
    return {"value": done_tris / wall / 1e6, "unit": "Mtri/s", "cores": threads, "kind": "port",
            "sample": f"{n_done} randomly chosen meshes of the scene ({done_tris} tris) built by the C "
                      f"restatement of marigold bvh, one mesh per task on {threads} threads, "
                      f"{wall:.1f} s wall; AABB setup included, TLAS excluded",
            "host": host_cores()}


def single_mesh(ctx, tuning: dict, K: int = 10):
    """The 1024x512-cell heightfield (1,048,576 triangles) built alone with the
    plain method and the given tuning (the library's defaults: the scene's
    autotuned settings are chosen for 2,000 meshes at once): level by level
    (the host reads each level's counts) and as a one-mesh scene plan (one HIP
    graph, no instances), with both trees compared.  The context's tuning is
    restored afterwards."""
    from marigold_amd import _native as N
    from marigold_amd import scenes
    from marigold_amd.scene import GpuScene
    saved = ctx.get_tuning()
    ctx.set_tuning(**tuning)
    pos, tris = scenes.heightfield(1024, 512, seed=2)
    n, nv = tris.shape[0], pos.shape[0]
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    dn, dl = ctx.alloc(32 * (2 * n - 1)), ctx.alloc(4 * (2 * n - 1))
    dp, dt = ctx.alloc(4 * n), ctx.alloc(12 * n)
    st = N.default_settings()
    for _ in range(3):
        nn = ctx.build_tris_device(d_pos, 12, nv, d_tris, n, st, dn, dl, dp, dt)
    ctx.synchronize()
    terms = tree_terms(dn.download(N.NODE_DTYPE, nn), dl.download(np.uint32, nn))
    t0 = time.perf_counter()
    for _ in range(K):
        ctx.build_tris_device(d_pos, 12, nv, d_tris, n, st, dn, dl, dp, dt)
    ctx.synchronize()
    lbl_ms = (time.perf_counter() - t0) / K * 1e3
    ref = dn.download(np.uint32, 8 * (2 * n - 1))
    g = GpuScene(ctx, [(pos, tris)], np.zeros((0, 4, 4), np.float32), np.zeros(0, np.uint32),
                 build=False)
    g.record_graph()
    for _ in range(3):
        g.build()
    dev = []
    t0 = time.perf_counter()
    for _ in range(K):
        ms, _, _ = g.plan.launch()
        dev.append(ms)
    wall_ms = (time.perf_counter() - t0) / K * 1e3
    same = bool(np.array_equal(g.download("blas_nodes", np.uint32, 8 * (2 * n - 1)), ref))
    g.free()
    for b in (d_pos, d_tris, dn, dl, dp, dt):
        b.free()
    ctx.set_tuning(**saved)
    dev_ms = float(np.median(dev))
    model = survey_bytes(n, nv, terms["s_large"], terms["s_small"])
    return {"workload": "1024x512-cell heightfield, 1,048,576 tris, plain method",
            "tuning": tuning, "graph_device_ms": dev_ms, "graph_wall_ms": wall_ms, "level_by_level_wall_ms": lbl_ms,
            "mtri_per_s_graph": n / (dev_ms * 1e-3) / 1e6,
            "byte_model_gb": model / 1e9,
            "byte_model_frac_of_hbm": model / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "trees_equal": same}


def _hits_of(ctx, sc_or_trace, rays_dev, n):
    h = ctx.alloc(16 * n)
    sc_or_trace(rays_dev, n, h)
    out = h.download(N_HIT(), n)
    h.free()
    return out


def N_HIT():
    from marigold_amd import _native as N
    return N.HIT_DTYPE


def same_hits(a, b) -> bool:
    return bool(np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
                and np.array_equal(a["prim"], b["prim"]) and np.array_equal(a["inst"], b["inst"]))


def heightfield_16m_rays(ctx, learned: dict):
    """The full-form check: the 1024x512-cell heightfield built by the plain
    and by the learned method, 8,388,608 primary rays (pinhole 4096 x 2048) and
    8,388,608 cosine bounces from the primary hits, closest hits compared."""
    from marigold_amd import _native as N
    from marigold_amd import scenes
    from marigold_amd.scene import look_at_camera
    pos, tris = scenes.heightfield(1024, 512, seed=2)
    n = tris.shape[0]
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    trees = []
    for method, extra in ((N.METHOD_PLAIN, {}), (N.METHOD_LEARNED, learned)):
        dn, dl = ctx.alloc(32 * (2 * n - 1)), ctx.alloc(4 * (2 * n - 1))
        dp, dt = ctx.alloc(4 * n), ctx.alloc(12 * n)
        ctx.build_tris_device(d_pos, 12, pos.shape[0], d_tris, n,
                              N.default_settings(method=method, **extra), dn, dl, dp, dt)
        trees.append((dn, dl, dp, dt))
    w, h = 4096, 2048
    nr = w * h
    cam = look_at_camera((0.5, 0.6, -0.2), (0.5, 0.0, 0.5), 0.6, w / h)
    r0, r1 = ctx.alloc(32 * nr), ctx.alloc(32 * nr)
    hp0, hl0, hp1, hl1 = (ctx.alloc(16 * nr) for _ in range(4))
    ctx.gen_primary(cam, w, h, r0)
    (pn, pl, pp, pt), (ln, ll, lp_, lt) = trees
    ctx.record(4)
    ctx.trace_closest(pn, pl, d_pos, 12, pt, pp, r0, nr, hp0)
    ctx.record(5)
    ctx.trace_closest(ln, ll, d_pos, 12, lt, lp_, r0, nr, hl0)
    ctx.gen_bounce(r0, hp0, nr, None, 0, None, 0, d_pos, 12, pt, 7, r1)
    ctx.record(6)
    ctx.trace_closest(pn, pl, d_pos, 12, pt, pp, r1, nr, hp1)
    ctx.record(7)
    ctx.trace_closest(ln, ll, d_pos, 12, lt, lp_, r1, nr, hl1)
    ms_primary, ms_bounce = ctx.elapsed_ms(4, 5), ctx.elapsed_ms(6, 7)
    a0, b0 = hp0.download(N.HIT_DTYPE, nr), hl0.download(N.HIT_DTYPE, nr)
    a1, b1 = hp1.download(N.HIT_DTYPE, nr), hl1.download(N.HIT_DTYPE, nr)
    # learned hits point at the learned tree's permuted list: compare by original index
    same = same_hits(a0, b0) and same_hits(a1, b1)
    out = {"rays": 2 * nr, "identical": same,
           "primary_hits": int((a0["prim"] != N.MISS).sum()),
           "bounce_hits": int((a1["prim"] != N.MISS).sum()),
           "trace_mrays_per_s": {"primary": nr / (ms_primary * 1e-3) / 1e6,
                                 "bounce": nr / (ms_bounce * 1e-3) / 1e6}}
    for b in (d_pos, d_tris, r0, r1, hp0, hl0, hp1, hl1):
        b.free()
    for t in trees:
        for b in t:
            b.free()
    return out


def scene_main(args):
    from marigold_amd import _native as N
    from marigold_amd import scenes
    from marigold_amd.autotune import autotune
    from marigold_amd.learned import DEFAULT_LEARNED, train_split_model
    from marigold_amd.scene import GpuScene, look_at_camera

    t0 = time.perf_counter()
    sc = scenes.batched_scene()
    meshes = list(zip(sc["positions"], sc["tris"]))
    n = int(sum(t.shape[0] for _, t in meshes))
    nv = int(sum(p.shape[0] for p, _ in meshes))
    log(f"scene: {len(meshes)} meshes, {n} tris, {nv} verts, {len(sc['instance_mesh'])} instances "
        f"(generated in {time.perf_counter() - t0:.1f} s)")
    ctx = N.Context(0)
    lib_tuning = ctx.get_tuning()  # the library's defaults, before any tuning
    model, train_info = train_split_model(ctx, epochs=(4, 2) if args.quick_train else (30, 15))
    ctx.set_model(model)
    log(f"split proposer trained on the GPU: {train_info}")
    scene_args = (ctx, meshes, sc["transforms"], sc["instance_mesh"])
    tuned = None
    learned = dict(DEFAULT_LEARNED)
    if args.no_autotune and args.medium_max is not None:
        ctx.set_tuning(medium_max_objects=args.medium_max)
    if not args.no_autotune:
        tuned = autotune(*scene_args, use_cache=False, log=log)
        if tuned["learned"]:
            learned = dict(tuned["learned"])
        log(f"autotuned: {tuned['tuning']} learned {learned}")

    # primary rays into the scene for the hit comparison of the two trees
    w, h = 1024, 1024
    cam = look_at_camera((0.0, 60.0, -260.0), (0.0, 0.0, 0.0), 0.9, 1.0)
    d_rays = ctx.alloc(32 * w * h)
    ctx.gen_primary(cam, w, h, d_rays)

    # ---- learned + batched (the headline)
    lea = GpuScene(*scene_args, mode="batched", build=False, method=N.METHOD_LEARNED,
                   learned=learned)
    # level by level, with every kernel class timed: which class dominates
    ctx.set_profiling(N.Context.KERNEL_KINDS)
    for _ in range(max(args.warmup, 1)):
        lea.build()
    prof = ctx.kernel_times(reset=True)
    dominant = max(prof, key=lambda k: prof[k][0])
    d = lea.d
    _, _, st = ctx.build_blas_batch(d["pos"], 12, nv, d["tris"], lea.mesh_desc, lea.blas_settings,
                                    d["blas_nodes"], d["blas_lens"], d["blas_perm"], d["tris_perm"])
    terms = terms_from_stats(st)
    # level by level, timed, with event pairs around the dominant class's
    # launches of the BLAS build (the TLAS build's launches are not counted)
    ctx.set_profiling([])
    ctx.synchronize()
    ctx.kernel_times(reset=True)
    ctx.record(0)
    for _ in range(args.steps):
        ctx.set_profiling([dominant])
        lea.build_blas()
        ctx.set_profiling([])
        lea.build_tlas()
    ctx.record(1)
    ctx.synchronize()
    ms_level = ctx.elapsed_ms(0, 1) / args.steps
    kt = ctx.kernel_times(reset=True)[dominant]
    dom_ms = kt[0] / kt[1] if kt[1] else 0.0  # per launch
    dom_launches = kt[1] / args.steps
    dom_bytes = kernel_bytes(dominant, n, nv, 12, terms) / max(dom_launches, 1)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    impl_bytes = sum(kernel_bytes(k, n, nv, 12, terms) for k in ("setup", "bin", "partition",
                                                                  "sweep", "small"))
    build_bytes = survey_bytes(n, nv, terms["s_large"], n)
    # the whole scene build as one graph replay: the timed region of `value`
    # (--no-graph: level-by-level builds instead, for profiler runs)
    if not args.no_graph:
        lea.record_graph()
        lea.build()
        ctx.synchronize()
        ctx.record(0)
        for _ in range(args.steps):
            lea.build()
        ctx.record(1)
        ctx.synchronize()
        ms_step = ctx.elapsed_ms(0, 1) / args.steps
    else:
        ms_step = ms_level
    sah_learned = lea.scene_sah()
    hits_learned = _hits_of(ctx, lea.trace, d_rays, w * h)
    log(f"learned+batched {ms_step:.2f} ms/scene ({'level by level' if args.no_graph else 'graph'}), "
        f"{ms_level:.2f} ms level by level, SAH {sah_learned:.4f}")
    lea.free()

    # ---- batched, plain split rule (bit-identical to the per-mesh plain builds)
    bat = GpuScene(*scene_args, mode="batched", build=False)
    if not args.no_graph:
        bat.record_graph()
    bat.build()
    ctx.synchronize()
    ctx.record(2)
    for _ in range(args.steps):
        bat.build()
    ctx.record(3)
    ctx.synchronize()
    ms_bat = ctx.elapsed_ms(2, 3) / args.steps
    sah_plain = bat.scene_sah()
    hits_plain = _hits_of(ctx, bat.trace, d_rays, w * h)
    # both traversals over the plain trees, 1M primary rays each, event-timed
    trace_ms = {}
    d_hits = ctx.alloc(16 * w * h)
    for name, fn in (("tree_independent", bat.trace), ("shader_exact", bat.raytrace)):
        fn(d_rays, w * h, d_hits)
        ctx.synchronize()
        ctx.record(2)
        for _ in range(3):
            fn(d_rays, w * h, d_hits)
        ctx.record(3)
        ctx.synchronize()
        trace_ms[name] = ctx.elapsed_ms(2, 3) / 3
    hits_shader = d_hits.download(N.HIT_DTYPE, w * h)
    d_hits.free()
    # every BLAS of the scene collapsed to the compressed 8-wide format in one call
    d_wide, d_tri = ctx.alloc(N.WIDE_NODE_BYTES * n), ctx.alloc(4 * n)
    wide_ms = []
    for _ in range(2):
        t_w = time.perf_counter()
        _, _, n_wide = ctx.wide8_collapse_batch(bat.d["blas_nodes"], None, bat.n_slots, bat.blas,
                                                bat.blas_node_counts, d_wide, n, d_tri, n)
        wide_ms.append((time.perf_counter() - t_w) * 1e3)
    d_wide.free()
    d_tri.free()
    log(f"8-wide collapse of all {len(meshes)} BLAS: {n_wide} wide nodes, {wide_ms[-1]:.1f} ms "
        f"(first call {wide_ms[0]:.1f} ms, scratch allocation included)")
    shader_same = float(np.mean((hits_shader["prim"] == hits_plain["prim"])
                                & (hits_shader["inst"] == hits_plain["inst"])))
    log(f"traversal of {w * h} rays: " + ", ".join(f"{k} {v:.2f} ms" for k, v in trace_ms.items())
        + f"; shader-exact hits on the same triangle as the tree-independent ones: {shader_same:.5f}")
    bat_nodes = bat.download("blas_nodes", np.uint32, 8 * bat.n_slots)
    log(f"batched plain {ms_bat:.2f} ms/scene ({'level by level' if args.no_graph else 'graph'}), "
        f"SAH {sah_plain:.4f}")
    bat.free()

    # ---- plain: one device build call per mesh (the reference's API shape)
    ms_plain, same_nodes = None, None
    if args.plain_steps > 0:
        pla = GpuScene(*scene_args, mode="plain", build=False)
        pla.build()
        ctx.synchronize()
        ctx.record(2)
        for _ in range(args.plain_steps):
            pla.build()
        ctx.record(3)
        ctx.synchronize()
        ms_plain = ctx.elapsed_ms(2, 3) / args.plain_steps
        same_nodes = bool(np.array_equal(bat_nodes,
                                         pla.download("blas_nodes", np.uint32, 8 * pla.n_slots)))
        # the same per-mesh builds once more with every kernel timed: how much of
        # the wall time is device work (the rest is launches and the host
        # reading each level's counts, per mesh)
        ctx.set_profiling(N.Context.KERNEL_KINDS)
        ctx.kernel_times(reset=True)
        pla.build()
        plain_kernel_ms = sum(v[0] for v in ctx.kernel_times(reset=True).values())
        ctx.set_profiling([])
        log(f"plain per mesh {ms_plain:.2f} ms/scene, of which kernels {plain_kernel_ms:.2f} ms")
        pla.free()
    d_rays.free()
    del bat_nodes

    full = None
    if not args.no_rays:
        log("16M-ray hit comparison on the heightfield")
        full = heightfield_16m_rays(ctx, learned)
    log("single mesh: the 1M-triangle heightfield alone, level by level and as a graph")
    single = single_mesh(ctx, lib_tuning)
    log(f"single mesh: graph {single['graph_device_ms']:.3f} ms device, "
        f"{single['graph_wall_ms']:.3f} ms wall; level by level {single['level_by_level_wall_ms']:.3f} ms")
    if args.calibrate_copy:
        # known-size device copies (256 MiB, then 1 GiB) for scaling the
        # FETCH_SIZE / WRITE_SIZE counters
        for nbytes in (1 << 28, 1 << 30):
            a, b = ctx.alloc(nbytes), ctx.alloc(nbytes)
            N.check(N.lib().mgbvh_copy_d2d(ctx.handle, b.ptr, a.ptr, nbytes))
            ctx.synchronize()
            a.free()
            b.free()
    model.close()
    out = {
        "metric": "BVH build Mtri/s, plain vs learned+batched, at SAH <=1.03x plain; 1 MI355X",
        "value": n / (ms_step * 1e-3) / 1e6,
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded)",
        "config": {"workload": "2,000-mesh scene (log-uniform 100..200,000 tris, 50/50 soup / "
                               "heightfield) + TLAS over 10,000 instances; whole scene build "
                               "replayed as one HIP graph",
                   "timed": "level-by-level builds" if args.no_graph else "graph replays",
                   "tris": n, "verts": nv, "meshes": len(meshes),
                   "instances": len(sc["instance_mesh"]), "max_depth": 32,
                   "method": "learned + batched", "learned": learned,
                   "tuning": ctx.get_tuning()},
        "level_by_level_ms": ms_level,
        "plain_per_mesh": None if ms_plain is None else
        {"value": n / (ms_plain * 1e-3) / 1e6, "ms_per_step": ms_plain, "steps": args.plain_steps,
         "kernel_ms_per_step": plain_kernel_ms,
         "speedup_of_value_over_kernel_time": plain_kernel_ms / ms_step,
         "node_arrays_equal_batched_plain": same_nodes},
        "batched_plain": {"value": n / (ms_bat * 1e-3) / 1e6, "ms_per_step": ms_bat},
        "speedup_vs_plain": None if ms_plain is None else ms_plain / ms_step,
        "sah": {"plain": sah_plain, "learned": sah_learned, "ratio": sah_learned / sah_plain},
        "learned_stats": {"window_splits": st["n_window_splits"],
                          "guard_fallbacks": st["n_guard_fallbacks"], "training": train_info},
        "autotune": None if tuned is None else {k: v for k, v in tuned.items() if k != "trials"},
        "scene_hits_identical": {"rays": w * h, "identical": same_hits(hits_plain, hits_learned),
                                 "hits": int((hits_plain["prim"] != N.MISS).sum())},
        "heightfield_16m_rays": full,
        "single_mesh": single,
        "wide8_collapse": {"blas": len(meshes), "wide_nodes": n_wide, "ms": wide_ms[-1],
                           "first_call_ms": wide_ms[0], "timed": "host wall time of one "
                           "mgbvh_wide8_collapse_batch_device call (it synchronises)"},
        "scene_traversal": {"rays": w * h, "trees": "batched plain",
                            "mrays_per_s": {k: w * h / (v * 1e-3) / 1e6 for k, v in trace_ms.items()},
                            "ms": trace_ms, "shader_exact_same_triangle_frac": shader_same},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dominant),
                     "kernel": dominant, "kernel_ms_per_launch": dom_ms,
                     "issue": measured_valu(dominant),
                     "launches_per_build": dom_launches, "alg_bytes_per_launch": dom_bytes,
                     "timed": "level-by-level builds of the same scene, HIP event pairs around "
                              "each launch of the kernel class in the BLAS build, on the build "
                              "stream"},
        "build_alg_gbs": build_bytes / (ms_step * 1e-3) / 1e9,
        "build_alg_bytes": build_bytes,
        "build_alg_model": "12N + 12V + 24N + 32 s_large + 28 N + 64N + 24N over the BLASes "
                           "(s_large: objects of the > 32-object internal nodes)",
        "build_impl_gbs": impl_bytes / (ms_step * 1e-3) / 1e9,
        "build_impl_bytes": impl_bytes,
        "kernel_ms_per_build": {k: v[0] / max(args.warmup, 1) for k, v in prof.items()},
        "work": terms,
    }
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline_scene(meshes)
        # RAYON_NUM_THREADS=1, the reference's own measurement setting (README.md:62-66)
        one = cpu_baseline_scene(meshes, budget_s=10.0, threads=1)
        out["cpu_baseline"]["one_thread"] = {"value": one["value"], "unit": one["unit"],
                                             "sample": one["sample"]}
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


def scene_plain_plan(args):
    """--plain-plan: the batched plain-rule scene build recorded as one scene
    plan (HIP graph) and replayed `steps` times (with --no-graph: level by
    level, every kernel class event-timed), for profiler runs of the timed
    region's kernels; no network is trained and no learned build runs."""
    from marigold_amd import _native as N
    from marigold_amd import scenes
    from marigold_amd.scene import GpuScene

    sc = scenes.batched_scene()
    meshes = list(zip(sc["positions"], sc["tris"]))
    n = int(sum(t.shape[0] for _, t in meshes))
    ctx = N.Context(0)
    if args.medium_max is not None:
        ctx.set_tuning(medium_max_objects=args.medium_max)
    g = GpuScene(ctx, meshes, sc["transforms"], sc["instance_mesh"], mode="batched", build=False)
    per_kernel = None
    if args.no_graph:
        ctx.set_profiling(N.Context.KERNEL_KINDS)
        for _ in range(args.warmup):
            g.build()
        ctx.kernel_times(reset=True)
    else:
        g.record_graph()
        for _ in range(args.warmup):
            g.build()
    ctx.synchronize()
    ctx.record(0)
    for _ in range(args.steps):
        g.build()
    ctx.record(1)
    ctx.synchronize()
    ms = ctx.elapsed_ms(0, 1) / args.steps
    if args.no_graph:
        per_kernel = {k: v[0] / args.steps for k, v in ctx.kernel_times(reset=True).items()}
        ctx.set_profiling([])
    wide8 = None
    if args.wide8:
        # every BLAS collapsed to the compressed 8-wide format in one call, 3 times
        d_wide, d_tri = ctx.alloc(N.WIDE_NODE_BYTES * n), ctx.alloc(4 * n)
        ms_w = []
        for _ in range(3):
            t_w = time.perf_counter()
            _, _, n_wide = ctx.wide8_collapse_batch(g.d["blas_nodes"], None, g.n_slots, g.blas,
                                                    g.blas_node_counts, d_wide, n, d_tri, n)
            ms_w.append((time.perf_counter() - t_w) * 1e3)
        d_wide.free()
        d_tri.free()
        wide8 = {"wide_nodes": n_wide, "ms_per_call": ms_w}
    print(json.dumps({"workload": "2,000-mesh scene, batched plain split rule",
                      "timed": "level-by-level builds" if args.no_graph else "scene-plan graph replays",
                      "tris": n, "steps": args.steps, "ms_per_step": ms,
                      "mtri_per_s": n / (ms * 1e-3) / 1e6, "tuning": ctx.get_tuning(),
                      "kernel_ms_per_build": per_kernel, "wide8_collapse": wide8}), flush=True)
    g.free()
    ctx.close()


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="scene", choices=["scene", "heightfield", "soup"])
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rays", action="store_true", help="skip the 16M-ray hit comparison")
    ap.add_argument("--no-autotune", action="store_true", help="default tuning and learned settings")
    ap.add_argument("--no-graph", action="store_true",
                    help="time level-by-level builds instead of graph replays (profiler runs)")
    ap.add_argument("--quick-train", action="store_true",
                    help="6 training epochs instead of 45 (profiling runs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="end with one 1 GiB device copy (counter calibration)")
    ap.add_argument("--plain-plan", action="store_true",
                    help="only the batched plain-rule scene plan, replayed (profiler runs of the "
                         "timed region)")
    ap.add_argument("--medium-max", type=int, default=None,
                    help="tuning: medium_max_objects of the --plain-plan build, or of "
                         "the scene builds with --no-autotune")
    ap.add_argument("--wide8", action="store_true",
                    help="--plain-plan: then collapse every BLAS to the 8-wide format 3 times")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: multi-GPU sharding "
              f"of the meshes is a design note only", file=sys.stderr)
        raise SystemExit(2)
    if args.plain_plan:
        return scene_plain_plan(args)
    if args.workload == "scene":
        return scene_main(args)

    from marigold_amd import _native as N

    pos, tris, desc = workload(args.workload)
    n, nv = tris.shape[0], pos.shape[0]
    print(f"[bench] workload {args.workload}: {n} tris, {nv} verts", file=sys.stderr, flush=True)
    ctx = N.Context(0)
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    d_nodes = ctx.alloc((2 * n - 1) * 32)
    d_lens = ctx.alloc((2 * n - 1) * 4)
    d_perm = ctx.alloc(n * 4)
    d_tout = ctx.alloc(n * 12)
    settings = N.default_settings(max_depth=32)
    nn = C.c_uint64()
    L = N.lib()

    def step():
        N.check(L.mgbvh_build_tris_device(ctx.handle, d_pos.ptr, 12, nv, d_tris.ptr, n, C.byref(settings),
                                          d_nodes.ptr, d_lens.ptr, d_perm.ptr, d_tout.ptr, None,
                                          C.byref(nn), None))

    # untimed: per-kernel profile to find the dominant kernel class
    ctx.set_profiling(N.Context.KERNEL_KINDS)
    for _ in range(max(args.warmup, 1)):
        step()
    prof = ctx.kernel_times(reset=True)
    dominant = max(prof, key=lambda k: prof[k][0])
    nodes = d_nodes.download(N.NODE_DTYPE, nn.value)
    lens = d_lens.download(np.uint32, nn.value)
    terms = tree_terms(nodes, lens)

    # timed: K builds, only the dominant kernel class carries event pairs
    ctx.set_profiling([dominant])
    ctx.synchronize()
    ctx.record(0)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.record(1)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ms_total = ctx.elapsed_ms(0, 1)
    ms_step = ms_total / args.steps
    kt = ctx.kernel_times(reset=True)[dominant]
    dom_ms = kt[0] / args.steps  # per build
    dom_bytes = kernel_bytes(dominant, n, nv, 12, terms)
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    build_bytes = sum(kernel_bytes(k, n, nv, 12, terms) for k in ("setup", "bin", "partition", "sweep",
                                                                   "small"))
    out = {
        "metric": "BVH build Mtri/s, plain vs learned+batched, at SAH <=1.03x plain; 1 MI355X",
        "value": n / (ms_step * 1e-3) / 1e6,
        "unit": "Mtri/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded)",
        "config": {"workload": desc, "tris": n, "verts": nv, "max_depth": 32, "method": "plain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dominant,
                     "kernel_ms_per_build": dom_ms, "alg_bytes_per_build": dom_bytes},
        "build_alg_gbs": build_bytes / (ms_step * 1e-3) / 1e9,
        "kernel_ms_per_build": {k: v[0] / max(args.warmup, 1) for k, v in prof.items()},
        "wall_s": wall,
    }
    ctx.close()
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(pos, tris)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
