"""Batched BLAS build against one-mesh-at-a-time builds, through the C ABI.

Every mesh of a batch must get exactly the tree, permutation, permuted
triangle list and split records of its own single-mesh build (which the
builder tests pin bit for bit to the oracle); a few meshes are also checked
against the oracle directly.
"""
import numpy as np
import pytest

from marigold_amd import _native as N
from marigold_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mixed_meshes():
    out = []
    for k, n in enumerate([1, 2, 31, 32, 33, 100, 2047, 2049, 5000, 0, 7, 300, 65, 4097]):
        out.append(scenes.soup(n, seed=500 + k) if n else
                   (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)))
    out.append(scenes.heightfield(40, 30, seed=7))
    out.append(scenes.clusters(3000, k=8, seed=8))
    # identical triangles: the median fallback inside a batch
    p = np.tile(np.array([[0.1, 0.2, 0.3], [0.4, 0.2, 0.3], [0.1, 0.6, 0.9]], np.float32), (90, 1))
    out.append((p, np.arange(270, dtype=np.uint32).reshape(90, 3)))
    out.append(scenes.heightfield(8, 3, seed=9))
    return out


def _concat(meshes):
    pos = np.ascontiguousarray(np.concatenate([m[0] for m in meshes]).astype(np.float32))
    tris = np.ascontiguousarray(np.concatenate([m[1] for m in meshes]).astype(np.uint32))
    desc = np.zeros(len(meshes), dtype=N.MESH_DTYPE)
    to = vo = 0
    for i, (p, t) in enumerate(meshes):
        desc[i] = (to, t.shape[0], vo, p.shape[0])
        to += t.shape[0]
        vo += p.shape[0]
    return pos, tris, desc


def _batch(ctx, meshes, max_depth=32, mesh_bounds=None):
    pos, tris, desc = _concat(meshes)
    n = tris.shape[0]
    slots = int(sum(2 * m[1].shape[0] - 1 for m in meshes if m[1].shape[0]))
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    d_nodes, d_lens = ctx.alloc(32 * slots), ctx.alloc(4 * slots)
    d_splits = ctx.alloc(16 * slots)
    d_perm, d_tout = ctx.alloc(4 * n), ctx.alloc(12 * n)
    blas, counts, st = ctx.build_blas_batch(d_pos, 12, pos.shape[0], d_tris, desc,
                                            N.default_settings(max_depth=max_depth), d_nodes,
                                            d_lens, d_perm, d_tout, d_splits, mesh_bounds)
    res = dict(blas=blas, counts=counts, stats=st,
               nodes=d_nodes.download(N.NODE_DTYPE, slots), lens=d_lens.download(np.uint32, slots),
               splits=d_splits.download(N.SPLIT_DTYPE, slots),
               perm=d_perm.download(np.uint32, n), tris=d_tout.download(np.uint32, 3 * n).reshape(-1, 3))
    for b in (d_pos, d_tris, d_nodes, d_lens, d_splits, d_perm, d_tout):
        b.free()
    return res, desc


def _check_against_single(ctx, meshes, res, desc, max_depth=32, mesh_bounds=None, tag=""):
    for i, (p, t) in enumerate(meshes):
        n = t.shape[0]
        assert res["blas"][i]["tri_offset"] == desc[i]["tri_offset"]
        if n == 0:
            assert res["counts"][i] == 0
            continue
        bounds = None if mesh_bounds is None else (mesh_bounds[i, :3], mesh_bounds[i, 3:])
        tt = t.copy()
        nodes, lens, perm, splits, _ = ctx.build_tris(
            p, tt, N.default_settings(max_depth=max_depth, bounds=bounds), want_splits=True)
        k = nodes.shape[0]
        nb = int(res["blas"][i]["node_offset"])
        to = int(desc[i]["tri_offset"])
        assert res["counts"][i] == k, (tag, i)
        got = res["nodes"][nb:nb + k].view(np.uint32).reshape(-1, 8)
        want = nodes.view(np.uint32).reshape(-1, 8)
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert bad.size == 0, (tag, "mesh", i, "n", n, "node", bad[:5], got[bad[:2]], want[bad[:2]])
        assert np.array_equal(res["lens"][nb:nb + k], lens), (tag, i)
        assert np.array_equal(res["splits"][nb:nb + k].view(np.uint32), splits.view(np.uint32)), (tag, i)
        assert np.array_equal(res["perm"][to:to + n], perm), (tag, i)
        assert np.array_equal(res["tris"][to:to + n], tt), (tag, i)


def test_batch_equals_single_builds(ctx):
    print("[case] batch of 18 mixed meshes", flush=True)
    meshes = _mixed_meshes()
    res, desc = _batch(ctx, meshes)
    _check_against_single(ctx, meshes, res, desc, tag="mixed")
    assert res["stats"]["total_object_count"] == sum(m[1].shape[0] for m in meshes)
    assert res["stats"]["n_median"] > 0


def test_batch_against_oracle(ctx):
    print("[case] batch vs oracle", flush=True)
    meshes = [scenes.soup(3000, seed=61), scenes.heightfield(30, 20, seed=62), scenes.soup(40, seed=63)]
    res, desc = _batch(ctx, meshes)
    for i, (p, t) in enumerate(meshes):
        ref = O.build(O.tri_aabbs(p, t))
        want = O.encode(ref["nodes"])
        nb, k = int(res["blas"][i]["node_offset"]), want.shape[0]
        assert np.array_equal(res["nodes"][nb:nb + k].view(np.uint32).reshape(-1, 8), want), i
        to, n = int(desc[i]["tri_offset"]), t.shape[0]
        assert np.array_equal(res["perm"][to:to + n], ref["list"]), i


@pytest.mark.parametrize("max_depth", [0, 2, 6])
def test_batch_depth_cut(ctx, max_depth):
    # cut leaves leave unused slots that are squeezed out inside each mesh
    print(f"[case] batch max_depth={max_depth}", flush=True)
    meshes = [scenes.soup(n, seed=700 + n) for n in (500, 3, 40, 0, 2500, 90)]
    res, desc = _batch(ctx, meshes, max_depth=max_depth)
    _check_against_single(ctx, meshes, res, desc, max_depth=max_depth, tag=f"cut{max_depth}")
    if max_depth < 6:
        assert res["stats"]["n_cutoff_leaves"] > 0
    # the sweep's one-node-per-lane mode (crowded levels) cuts the same leaves
    old = ctx.get_tuning()
    try:
        ctx.set_tuning(sweep_pack_min_nodes=1)
        lanes, _ = _batch(ctx, meshes, max_depth=max_depth)
    finally:
        ctx.set_tuning(**{k: v for k, v in old.items() if k != "compute_units"})
    # slots past a mesh's node count (the squeezed-out ones) are written by no build
    valid = np.zeros(res["nodes"].shape[0], bool)
    for i in range(len(meshes)):
        nb = int(res["blas"][i]["node_offset"])
        valid[nb:nb + int(res["counts"][i])] = True
    assert np.array_equal(lanes["counts"], res["counts"])
    for k in ("nodes", "lens", "splits", "perm", "tris"):
        a, b = np.asarray(lanes[k]), np.asarray(res[k])
        if k in ("nodes", "lens", "splits"):
            a, b = a[valid], b[valid]
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k


def test_batch_mesh_bounds(ctx):
    print("[case] batch with per-mesh bounds", flush=True)
    meshes = [scenes.soup(n, seed=800 + n) for n in (1000, 50, 20)]
    mb = np.array([np.concatenate([p.min(axis=0) - 0.25, p.max(axis=0) + 0.5]) for p, _ in meshes],
                  np.float32)
    res, desc = _batch(ctx, meshes, mesh_bounds=mb)
    _check_against_single(ctx, meshes, res, desc, mesh_bounds=mb, tag="bounds")


def test_batch_errors(ctx):
    meshes = [scenes.soup(10, seed=1), scenes.soup(20, seed=2)]
    pos, tris, desc = _concat(meshes)
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    d_nodes, d_lens, d_perm = ctx.alloc(32 * 60), ctx.alloc(4 * 60), ctx.alloc(4 * 30)
    s = N.default_settings()
    bad = desc.copy()
    bad[1]["tri_offset"] = 11  # not contiguous
    with pytest.raises(N.MgbvhError) as e:
        ctx.build_blas_batch(d_pos, 12, pos.shape[0], d_tris, bad, s, d_nodes, d_lens, d_perm)
    assert e.value.code == N.ERR_INVALID_ARG
    bad = desc.copy()
    bad[1]["vertex_count"] = 10_000  # beyond the vertex array
    with pytest.raises(N.MgbvhError) as e:
        ctx.build_blas_batch(d_pos, 12, pos.shape[0], d_tris, bad, s, d_nodes, d_lens, d_perm)
    assert e.value.code == N.ERR_INDEX
    bad = desc.copy()
    bad[0]["vertex_count"] = 5  # triangle indices of mesh 0 reach 29
    with pytest.raises(N.MgbvhError) as e:
        ctx.build_blas_batch(d_pos, 12, pos.shape[0], d_tris, bad, s, d_nodes, d_lens, d_perm)
    assert e.value.code == N.ERR_INDEX
    for b in (d_pos, d_tris, d_nodes, d_lens, d_perm):
        b.free()


def test_scene_plain_and_batched_identical(ctx):
    # a 60-mesh scene with 300 instances: the batched build, the GPU instance
    # boxes and the TLAS equal the per-mesh plain builds and the oracle's boxes
    print("[case] 60-mesh scene plain vs batched", flush=True)
    from marigold_amd.scene import GpuScene
    sc = scenes.batched_scene(n_meshes=60, n_instances=300, lo=20, hi=20_000, seed=11)
    meshes = list(zip(sc["positions"], sc["tris"]))
    a = GpuScene(ctx, meshes, sc["transforms"], sc["instance_mesh"], mode="plain")
    b = GpuScene(ctx, meshes, sc["transforms"], sc["instance_mesh"], mode="batched")
    for name, dt, cnt in (("blas_nodes", np.uint32, 8 * a.n_slots), ("blas_lens", np.uint32, a.n_slots),
                          ("blas_perm", np.uint32, a.n_tris), ("tris_perm", np.uint32, 3 * a.n_tris),
                          ("boxes", np.float32, 6 * a.n_instances),
                          ("tlas_nodes", np.uint32, 8 * a.n_tlas_nodes),
                          ("tlas_perm", np.uint32, a.n_instances)):
        x, y = a.download(name, dt, cnt), b.download(name, dt, cnt)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), name
    assert np.array_equal(a.blas_node_counts, b.blas_node_counts)
    assert a.n_tlas_nodes == b.n_tlas_nodes == 2 * a.n_instances - 1
    # instance boxes against the oracle (value equality: -0 and +0 may differ)
    nodes = a.download("blas_nodes", N.NODE_DTYPE, a.n_slots)
    roots = nodes[a.node_offsets]
    want = O.instance_boxes(roots["min"][sc["instance_mesh"]], roots["max"][sc["instance_mesh"]], a.o2w)
    got = a.download("boxes", np.float32, 6 * a.n_instances).reshape(-1, 6)
    assert np.array_equal(got, want)
    # the TLAS against the oracle over those boxes
    ref = O.build(want)
    tl = a.download("tlas_nodes", np.uint32, 8 * a.n_tlas_nodes).reshape(-1, 8)
    assert np.array_equal(tl, O.encode(ref["nodes"]))
    a.free()
    b.free()


def _scene_arrays(sc, n):
    """The scene's output arrays.  Node and length slots past a BLAS's node
    count (left by depth-cut leaves, whose unused slots the compaction drops)
    are written by no build, so they are zeroed here before comparing."""
    out = {name: sc.download(name, dt, cnt(sc)) for name, dt, cnt in (
        ("blas_nodes", np.uint32, lambda s: 8 * s.n_slots), ("blas_lens", np.uint32, lambda s: s.n_slots),
        ("blas_perm", np.uint32, lambda s: s.n_tris), ("tris_perm", np.uint32, lambda s: 3 * s.n_tris),
        ("boxes", np.float32, lambda s: 6 * s.n_instances),
        ("tlas_nodes", np.uint32, lambda s: 8 * (2 * s.n_instances - 1)),
        ("tlas_perm", np.uint32, lambda s: s.n_instances))}
    valid = np.zeros(sc.n_slots, bool)
    for off, cnt in zip(sc.node_offsets, sc.blas_node_counts):
        valid[int(off):int(off) + int(cnt)] = True
    out["blas_nodes"].reshape(-1, 8)[~valid] = 0
    out["blas_lens"][~valid] = 0
    return out


def test_scene_graph_replays_equal_level_by_level(ctx):
    # the whole scene build recorded as one HIP graph (persistent level kernels,
    # fixed level count) gives the level-by-level batched build's arrays, also
    # on a replay after the vertices moved
    print("[case] scene graph plan vs level-by-level", flush=True)
    from marigold_amd.scene import GpuScene
    sc = scenes.batched_scene(n_meshes=40, n_instances=200, lo=20, hi=30_000, seed=12)
    meshes = list(zip(sc["positions"], sc["tris"]))
    ref = GpuScene(ctx, meshes, sc["transforms"], sc["instance_mesh"], mode="batched")
    want = _scene_arrays(ref, None)
    g = GpuScene(ctx, meshes, sc["transforms"], sc["instance_mesh"], mode="batched", build=False)
    g.record_graph()
    for rep in range(2):
        g.build()
        got = _scene_arrays(g, None)
        for k in want:
            assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), (rep, k)
        assert np.array_equal(g.blas_node_counts, ref.blas_node_counts)
    # new vertex positions in the same buffers: the replay equals a fresh build
    moved = [(p * np.float32(1.5) + np.float32(0.25), t) for p, t in meshes]
    ref2 = GpuScene(ctx, moved, sc["transforms"], sc["instance_mesh"], mode="batched")
    g.d["pos"].upload(np.ascontiguousarray(np.concatenate([p for p, _ in moved])))
    g.build()
    got, want2 = _scene_arrays(g, None), _scene_arrays(ref2, None)
    for k in want2:
        assert np.array_equal(got[k].view(np.uint32), want2[k].view(np.uint32)), ("moved", k)
    assert g.plan.info()["launches_rebuilt"] == 0
    # outliers at 4, 16, ..., 4^24 on the largest mesh: its tree needs about 8
    # more levels than the graph records, so the launch rebuilds level by level
    big = int(np.argmax([t.shape[0] for _, t in meshes]))
    deep = [(p.copy(), t) for p, t in meshes]
    k = np.arange(24)
    q = deep[big][0]
    q[k * 7 % q.shape[0], 0] = np.float32(4.0) ** (k + 1).astype(np.float32)
    ref3 = GpuScene(ctx, deep, sc["transforms"], sc["instance_mesh"], mode="batched")
    g.d["pos"].upload(np.ascontiguousarray(np.concatenate([p for p, _ in deep])))
    g.build()
    assert g.plan.info()["launches_rebuilt"] == 1, g.plan.info()
    got, want3 = _scene_arrays(g, None), _scene_arrays(ref3, None)
    for k_ in want3:
        assert np.array_equal(got[k_].view(np.uint32), want3[k_].view(np.uint32)), ("deep", k_)
    assert np.array_equal(g.blas_node_counts, ref3.blas_node_counts)
    for s_ in (ref, g, ref2, ref3):
        s_.free()


@pytest.mark.parametrize("tuning", [dict(bin_blocks_per_cu=1, partition_blocks_per_cu=1,
                                         sweep_blocks_per_cu=1, small_blocks_per_cu=1,
                                         partition_waves=4, bin_waves=1,
                                         sweep_pack_min_nodes=1),
                                    dict(bin_blocks_per_cu=32, partition_blocks_per_cu=256,
                                         sweep_blocks_per_cu=4, small_blocks_per_cu=256,
                                         partition_waves=1, bin_waves=2,
                                         sweep_pack_min_nodes=1 << 30),
                                    dict(medium_max_objects=33),
                                    dict(medium_max_objects=512),
                                    dict(medium_max_objects=2048, partition_waves=4)])
def test_tuning_does_not_change_results(ctx, tuning):
    print(f"[case] tuning {tuning}", flush=True)
    meshes = _mixed_meshes()
    base, desc = _batch(ctx, meshes)
    old = ctx.get_tuning()
    ctx.set_tuning(**tuning)
    try:
        res, _ = _batch(ctx, meshes)
    finally:
        ctx.set_tuning(**{k: v for k, v in old.items() if k != "compute_units"})
    for k in ("nodes", "lens", "splits", "perm", "tris"):
        assert np.array_equal(np.asarray(res[k]).view(np.uint8), np.asarray(base[k]).view(np.uint8)), k


def test_tuning_limits(ctx):
    """Grid sizes 1..256 blocks per CU, 1 or 4 partition wavefronts; anything
    else is refused and leaves the tuning as it was."""
    print("[case] tuning limits", flush=True)
    old = ctx.get_tuning()
    try:
        ctx.set_tuning(bin_blocks_per_cu=256, small_blocks_per_cu=1)
        assert ctx.get_tuning()["bin_blocks_per_cu"] == 256
        for bad in (dict(bin_blocks_per_cu=257), dict(partition_blocks_per_cu=0),
                    dict(partition_waves=2), dict(bin_waves=5)):
            with pytest.raises(N.MgbvhError):
                ctx.set_tuning(**bad)
        assert ctx.get_tuning()["bin_blocks_per_cu"] == 256
    finally:
        ctx.set_tuning(**{k: v for k, v in old.items() if k != "compute_units"})


def test_small_windows_same_trees(ctx):
    """k_small ranks windows of subtrees by size only when even the fewest
    possible subtrees fill its grid; a one-block-per-CU grid makes a 320 k
    triangle batch take that path, and the trees equal the plain pairs'."""
    print("[case] size-ranked subtree windows vs plain pairs", flush=True)
    meshes = [scenes.soup(200000, seed=901), scenes.heightfield(300, 200, seed=902)]
    base, _ = _batch(ctx, meshes)
    old = ctx.get_tuning()
    try:
        ctx.set_tuning(small_blocks_per_cu=1)
        win, _ = _batch(ctx, meshes)
    finally:
        ctx.set_tuning(**{k: v for k, v in old.items() if k != "compute_units"})
    for k in ("nodes", "lens", "splits", "perm", "tris"):
        assert np.array_equal(np.asarray(win[k]).view(np.uint8), np.asarray(base[k]).view(np.uint8)), k


def test_learned_plan_keeps_its_model(ctx):
    """A learned scene plan builds with the model as it was when the plan was
    recorded: after set_params, and after set_model(None), its replays (and a
    replay that falls back to a level-by-level rebuild) give the trees of the
    original parameters."""
    print("[case] learned scene plan keeps its model", flush=True)
    from marigold_amd.scene import GpuScene
    sc = scenes.batched_scene(n_meshes=30, n_instances=100, lo=200, hi=20_000, seed=14)
    meshes = list(zip(sc["positions"], sc["tris"]))
    learned = dict(learned_min_objects=64, learned_window=1, learned_guard=10.0)
    model = N.Model(ctx, seed=5)
    p0 = model.params()
    ctx.set_model(model)
    args = (ctx, meshes, sc["transforms"], sc["instance_mesh"])
    kw = dict(mode="batched", method=N.METHOD_LEARNED, learned=learned)
    ref = GpuScene(*args, **kw)
    want = _scene_arrays(ref, None)
    # geometry whose trees need more levels than the plan records (redo path)
    big = int(np.argmax([t.shape[0] for _, t in meshes]))
    deep = [(p.copy(), t) for p, t in meshes]
    k = np.arange(24)
    q = deep[big][0]
    q[k * 7 % q.shape[0], 0] = np.float32(4.0) ** (k + 1).astype(np.float32)
    ref_deep = GpuScene(ctx, deep, sc["transforms"], sc["instance_mesh"], **kw)
    want_deep = _scene_arrays(ref_deep, None)
    g = GpuScene(*args, build=False, **kw)
    g.record_graph()
    g.build()
    got = _scene_arrays(g, None)
    for key in want:
        assert np.array_equal(got[key].view(np.uint32), want[key].view(np.uint32)), ("first", key)
    # other parameters change the context's trees ...
    rng = np.random.default_rng(3)
    model.set_params((p0 + rng.normal(0.0, 1.0, p0.shape)).astype(np.float32))
    other = GpuScene(*args, **kw)
    changed = _scene_arrays(other, None)
    assert not np.array_equal(changed["blas_nodes"].view(np.uint32),
                              want["blas_nodes"].view(np.uint32))
    # ... but not the plan's
    g.build()
    got = _scene_arrays(g, None)
    for key in want:
        assert np.array_equal(got[key].view(np.uint32), want[key].view(np.uint32)), ("params", key)
    # no model on the context: replays and a level-by-level redo still build
    ctx.set_model(None)
    g.d["pos"].upload(np.ascontiguousarray(np.concatenate([p for p, _ in deep])))
    g.build()
    assert g.plan.info()["launches_rebuilt"] == 1, g.plan.info()
    got = _scene_arrays(g, None)
    for key in want_deep:
        assert np.array_equal(got[key].view(np.uint32), want_deep[key].view(np.uint32)), ("deep", key)
    for s_ in (ref, ref_deep, g, other):
        s_.free()
    model.close()


def test_autotune_tries_every_candidate_and_keeps_results(ctx):
    # the autotuner times every grid and medium-node candidate on the scene's
    # graph, applies the fastest, and the trees it builds then are the ones an
    # untuned build gives (none of these settings changes results)
    print("[case] autotune on a 30-mesh scene", flush=True)
    from marigold_amd import autotune as A
    from marigold_amd.scene import GpuScene
    sc = scenes.batched_scene(n_meshes=30, n_instances=120, lo=20, hi=20_000, seed=14)
    meshes = list(zip(sc["positions"], sc["tris"]))
    old = ctx.get_tuning()
    model = ctx.model
    ctx.set_model(None)
    try:
        ref = GpuScene(ctx, meshes, sc["transforms"], sc["instance_mesh"], mode="batched")
        want = _scene_arrays(ref, None)
        ref.free()
        res = A.autotune(ctx, meshes, sc["transforms"], sc["instance_mesh"], reps=1,
                         use_cache=False)
        assert len(res["trials"]) == len(A.GRID_CANDIDATES) + len(A.MEDIUM_CANDIDATES)
        assert res["learned"] is None and res["ms"] > 0
        tuned = ctx.get_tuning()
        for k, v in res["tuning"].items():
            assert tuned[k] == v, (k, tuned[k], v)
        g = GpuScene(ctx, meshes, sc["transforms"], sc["instance_mesh"], mode="batched",
                     build=False)
        g.record_graph()
        g.build()
        got = _scene_arrays(g, None)
        g.free()
        for k in want:
            assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), k
    finally:
        ctx.set_tuning(**{k: v for k, v in old.items() if k != "compute_units"})
        ctx.set_model(model)
