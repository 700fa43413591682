"""The benchmark's workloads at their configured sizes, on the GPU.

* the 2,000-mesh scene (log-uniform 100..200,000 triangles, ~52M in all, 10,000
  instances) built as one recorded scene graph: every BLAS of a seeded sample of
  meshes (the 10 largest among them) and the TLAS equal the oracle's trees
  (marigold/src/app/data/scene/model.rs:76-95, geometry/mod.rs:175-199);
* 8,000,000 triangles in 64 Gaussian clusters built with the learned method at a
  dispatch that takes thousands of window splits: a valid tree, SAH within
  1.03x of the plain tree's, identical closest hits on 1M rays;
* the 1024x512-cell heightfield built with the learned method: 8,388,608
  pinhole primaries (4096 x 2048) and 8,388,608 cosine bounces hit exactly what
  they hit in the plain tree, and a seeded subset equals the brute force.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from marigold_amd import _native as N
from marigold_amd import scenes
from marigold_amd.learned import train_split_model
from marigold_amd.scene import GpuScene, look_at_camera
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model(ctx):
    m, info = train_split_model(ctx, epochs=(6, 3))
    print(f"[train] {info}", flush=True)
    yield m
    ctx.set_model(None)
    m.close()


def _same_hits(a, b):
    return (np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
            and np.array_equal(a["prim"], b["prim"]))


def test_full_scene_against_oracle(ctx):
    print("[case] 2,000-mesh scene (graph replay) vs the oracle on 200 meshes and the TLAS",
          flush=True)
    sc = scenes.batched_scene()
    meshes = list(zip(sc["positions"], sc["tris"]))
    g = GpuScene(ctx, meshes, sc["transforms"], sc["instance_mesh"], mode="batched", build=False)
    g.record_graph()
    g.build()
    assert g.n_tris > 50_000_000 and g.n_instances == 10_000
    counts = g.counts
    rng = np.random.default_rng(41)
    largest = np.argsort(counts)[-10:]
    sample = np.unique(np.concatenate([largest, rng.choice(len(meshes), 190, replace=False)]))
    assert len(sample) >= 200 - 10 and set(largest) <= set(sample)

    def one(m):
        p, t = meshes[m]
        r = O.build(O.tri_aabbs(p, t), max_depth=32)
        return m, O.encode(r["nodes"]), r["list"]

    with ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(one, sample))
    for m, want, lst in refs:
        n, nb, to = int(counts[m]), int(g.node_offsets[m]), int(g.tri_offsets[m])
        assert g.blas_node_counts[m] == 2 * n - 1, m
        got = g.d["blas_nodes"].download(np.uint32, 8 * (2 * n - 1), offset=8 * nb).reshape(-1, 8)
        assert np.array_equal(got, want), ("mesh", m, n)
        perm = g.d["blas_perm"].download(np.uint32, n, offset=to)
        assert np.array_equal(perm, lst), ("perm", m)
        tp = g.d["tris_perm"].download(np.uint32, 3 * n, offset=3 * to).reshape(-1, 3)
        assert np.array_equal(tp, meshes[m][1][lst]), ("triangles", m)
    # the TLAS over the instance boxes (from each mesh's root box)
    roots = np.stack([g.d["blas_nodes"].download(np.uint32, 8, offset=8 * int(nb))
                      for nb in g.node_offsets]).view(np.float32)
    mn, mx = roots[:, 0:3], roots[:, 4:7]
    boxes = O.instance_boxes(mn[sc["instance_mesh"]], mx[sc["instance_mesh"]], g.o2w)
    got_boxes = g.download("boxes", np.float32, 6 * g.n_instances).reshape(-1, 6)
    assert np.array_equal(got_boxes, boxes)
    ref = O.build(boxes, max_depth=32)
    tl = g.download("tlas_nodes", np.uint32, 8 * g.n_tlas_nodes).reshape(-1, 8)
    assert np.array_equal(tl, O.encode(ref["nodes"]))
    assert np.array_equal(g.download("tlas_perm", np.uint32, g.n_instances), ref["list"])
    g.free()


def _device_build(ctx, d_pos, nverts, d_tris, n, settings):
    slots = 2 * n - 1
    dn, dl, dp, dt = ctx.alloc(32 * slots), ctx.alloc(4 * slots), ctx.alloc(4 * n), ctx.alloc(12 * n)
    k = ctx.build_tris_device(d_pos, 12, nverts, d_tris, n, settings, dn, dl, dp, dt)
    return k, (dn, dl, dp, dt)


def test_learned_8m_clusters(ctx, model):
    print("[case] learned build of 8M triangles in 64 Gaussian clusters", flush=True)
    pos, tris = scenes.clusters(8_000_000, k=64, seed=4)
    n = tris.shape[0]
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    ctx.set_model(model)
    kp, plain = _device_build(ctx, d_pos, pos.shape[0], d_tris, n, N.default_settings())
    learned_kw = dict(learned_min_objects=1024, learned_window=3, learned_guard=0.02)
    s = N.default_settings(method=N.METHOD_LEARNED, **learned_kw)
    # window splits come from the builder's statistics (host-buffer entry point)
    t = tris.copy()
    nodes, lens, perm, _, st = ctx.build_tris(pos, t, s)
    print(f"window splits {st['n_window_splits']}, guard fallbacks {st['n_guard_fallbacks']}",
          flush=True)
    assert st["n_window_splits"] > 1000
    k = nodes.shape[0]
    assert kp <= 2 * n - 1 and k <= 2 * n - 1  # depth cuts (max_depth 32) may merge leaves
    # a valid tree: the leaves' ranges cover every object once, children
    # partition their parent, boxes are exact unions
    child = nodes["child_node"].astype(np.int64)
    start = nodes["start_index"].astype(np.int64)
    ln = lens.astype(np.int64)
    internal = np.nonzero(child)[0]
    leaf = np.nonzero(child == 0)[0]
    order = leaf[np.argsort(start[leaf])]
    assert start[order[0]] == 0 and np.all(start[order[1:]] == start[order[:-1]] + ln[order[:-1]])
    assert start[order[-1]] + ln[order[-1]] == n
    assert np.array_equal(np.sort(perm), np.arange(n))
    gt, lt = child[internal], child[internal] + 1
    assert np.array_equal(np.sort(np.concatenate([gt, lt])), np.arange(1, k))
    assert np.all(start[lt] == start[internal]) and np.all(start[gt] == start[internal] + ln[lt])
    assert np.all(ln[lt] + ln[gt] == ln[internal])
    ab = O.tri_aabbs(pos, tris)[perm]
    lo = np.minimum.reduceat(ab[:, :3], start[order])
    hi = np.maximum.reduceat(ab[:, 3:], start[order])
    assert np.array_equal(nodes["min"][order], lo) and np.array_equal(nodes["max"][order], hi)
    assert np.array_equal(nodes["min"][internal], np.minimum(nodes["min"][gt], nodes["min"][lt]))
    assert np.array_equal(nodes["max"][internal], np.maximum(nodes["max"][gt], nodes["max"][lt]))
    # SAH of both trees (the library's tree-SAH kernel)
    dn, dl = ctx.to_device(nodes), ctx.to_device(lens)
    blas = np.zeros(1, N.BLAS_DTYPE)
    sah_l = ctx.tree_sah(dn, dl, blas, np.array([k], np.uint32))[0]
    sah_p = ctx.tree_sah(plain[0], plain[1], blas, np.array([kp], np.uint32))[0]
    print(f"SAH plain {sah_p:.4f} learned {sah_l:.4f} ratio {sah_l / sah_p:.5f}", flush=True)
    assert sah_l <= 1.03 * sah_p
    # identical closest hits on 1,048,576 rays (camera rays into the clusters)
    w = h = 1024
    cam = look_at_camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), 0.9, 1.0)
    d_r = ctx.alloc(32 * w * h)
    ctx.gen_primary(cam, w, h, d_r)
    dtl, dpl = ctx.to_device(t), ctx.to_device(perm)
    h0, h1 = ctx.alloc(16 * w * h), ctx.alloc(16 * w * h)
    ctx.trace_closest(plain[0], plain[1], d_pos, 12, plain[3], plain[2], d_r, w * h, h0)
    ctx.trace_closest(dn, dl, d_pos, 12, dtl, dpl, d_r, w * h, h1)
    a, b = h0.download(N.HIT_DTYPE, w * h), h1.download(N.HIT_DTYPE, w * h)
    assert (a["prim"] != N.MISS).sum() > w * h // 20
    assert _same_hits(a, b)
    for buf in (d_pos, d_tris, d_r, dtl, dpl, h0, h1, dn, dl) + plain:
        buf.free()


def test_learned_heightfield_16m_rays(ctx, model):
    print("[case] learned heightfield 1024x512, 8.4M primary + 8.4M bounce rays", flush=True)
    pos, tris = scenes.heightfield(1024, 512, seed=2)
    n = tris.shape[0]
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    ctx.set_model(model)
    learned_kw = dict(learned_min_objects=1024, learned_window=3, learned_guard=0.02)
    _, pl = _device_build(ctx, d_pos, pos.shape[0], d_tris, n, N.default_settings())
    _, le = _device_build(ctx, d_pos, pos.shape[0], d_tris, n,
                          N.default_settings(method=N.METHOD_LEARNED, **learned_kw))
    # the learned tree differs from the plain one
    assert not np.array_equal(pl[2].download(np.uint32, n), le[2].download(np.uint32, n))
    w, h = 4096, 2048
    nr = w * h
    cam = look_at_camera((0.5, 0.6, -0.2), (0.5, 0.0, 0.5), 0.6, w / h)
    r0, r1 = ctx.alloc(32 * nr), ctx.alloc(32 * nr)
    hp0, hl0, hp1, hl1 = (ctx.alloc(16 * nr) for _ in range(4))
    ctx.gen_primary(cam, w, h, r0)
    ctx.trace_closest(pl[0], pl[1], d_pos, 12, pl[3], pl[2], r0, nr, hp0)
    ctx.trace_closest(le[0], le[1], d_pos, 12, le[3], le[2], r0, nr, hl0)
    ctx.gen_bounce(r0, hp0, nr, None, 0, None, 0, d_pos, 12, pl[3], 7, r1)
    ctx.trace_closest(pl[0], pl[1], d_pos, 12, pl[3], pl[2], r1, nr, hp1)
    ctx.trace_closest(le[0], le[1], d_pos, 12, le[3], le[2], r1, nr, hl1)
    a0, b0 = hp0.download(N.HIT_DTYPE, nr), hl0.download(N.HIT_DTYPE, nr)
    a1, b1 = hp1.download(N.HIT_DTYPE, nr), hl1.download(N.HIT_DTYPE, nr)
    assert (a0["prim"] != N.MISS).sum() > nr // 4 and (a1["prim"] != N.MISS).sum() > nr // 20
    assert _same_hits(a0, b0) and _same_hits(a1, b1)
    # a seeded subset of both ray sets against the brute force over all triangles
    rng = np.random.default_rng(17)
    sub = np.sort(rng.choice(nr, 96, replace=False))
    for tag, rays_dev, hits in (("primary", r0, a0), ("bounce", r1, a1)):
        rr = np.concatenate([rays_dev.download(N.RAY_DTYPE, 1, offset=int(i)) for i in sub])
        tb, pb = O.closest_hit_brute(pos, tris, np.concatenate([rr["org"], rr["dir"]], axis=1))
        assert np.array_equal(hits["t"][sub].view(np.uint32), tb.view(np.uint32)), tag
        assert np.array_equal(hits["prim"][sub], pb), tag
    for b in (d_pos, d_tris, r0, r1, hp0, hl0, hp1, hl1) + pl + le:
        b.free()
