"""Compressed 8-wide BVH on the GPU (mgbvh_wide8_collapse_device,
mgbvh_trace_wide8_device) against the oracle's collapse: byte-identical wide
nodes and object order, the same collapse cost, and traversal hits identical
to the binary tree's."""
import time

import numpy as np
import pytest

from marigold_amd import _native as N
from marigold_amd import scenes
from marigold_amd.scene import look_at_camera
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _collapse_both(ctx, pos, tris, max_depth=32, **params):
    t = tris.copy()
    nodes, lens, perm, _, _ = ctx.build_tris(pos, t, N.default_settings(max_depth=max_depth))
    k = nodes.shape[0]
    n = tris.shape[0]
    d_nodes, d_lens = ctx.to_device(nodes), ctx.to_device(lens)
    d_wide, d_tri = ctx.alloc(N.WIDE_NODE_BYTES * max(n - 1, 1)), ctx.alloc(4 * n)
    t0 = time.perf_counter()
    nw, cost = ctx.wide8_collapse(d_nodes, d_lens if max_depth < 32 else None, k, d_wide,
                                  max(n - 1, 1), d_tri, n, **params)
    ms = (time.perf_counter() - t0) * 1e3
    wide = d_wide.download(np.uint8, N.WIDE_NODE_BYTES * nw).reshape(nw, N.WIDE_NODE_BYTES)
    tri_index = d_tri.download(np.uint32, n)
    ref_w, ref_t, ref_c = O.wide8_collapse(nodes.view(np.uint32).reshape(-1, 8),
                                           lens if max_depth < 32 else None, **params)
    for b in (d_nodes, d_lens, d_wide, d_tri):
        b.free()
    return (wide, tri_index, cost), (ref_w, ref_t, ref_c), (nodes, lens, perm, t), ms


def _assert_same(tag, got, ref):
    wide, tri_index, cost = got
    rw, rt, rc = ref
    assert wide.shape == rw.shape, (tag, wide.shape, rw.shape)
    bad = np.flatnonzero((wide != rw).any(axis=1))
    assert bad.size == 0, (tag, bad[:5], wide[bad[:1]], rw[bad[:1]])
    assert np.array_equal(tri_index, rt), tag
    assert np.float32(cost).view(np.uint32) == np.float32(rc).view(np.uint32), (tag, cost, rc)


@pytest.mark.parametrize("kind", ["soup", "heightfield", "clusters", "one", "two", "three"])
def test_collapse_matches_oracle(ctx, kind):
    print(f"[case] wide8 collapse {kind}", flush=True)
    pos, tris = {"soup": lambda: scenes.soup(3000, seed=21),
                 "heightfield": lambda: scenes.heightfield(60, 40, seed=22),
                 "clusters": lambda: scenes.clusters(5000, k=7, seed=23),
                 "one": lambda: scenes.soup(1, seed=24, extent=0.5),
                 "two": lambda: scenes.soup(2, seed=25, extent=0.5),
                 "three": lambda: scenes.soup(3, seed=26, extent=0.5)}[kind]()
    got, ref, _, _ = _collapse_both(ctx, pos, tris)
    _assert_same(kind, got, ref)


@pytest.mark.parametrize("params", [dict(c_node=1.0, c_prim=2.0, p_max=3),
                                    dict(c_node=0.2, c_prim=0.3, p_max=1),
                                    dict(c_node=50.0, c_prim=0.3, p_max=2),
                                    dict(c_node=3e38, c_prim=0.3, p_max=3)])
def test_collapse_parameters(ctx, params):
    # c_node = 3e38 overflows the cost sums to +inf on the scene scaled by 4:
    # subtrees of more than p_max objects must still become internal nodes
    print(f"[case] wide8 collapse parameters {params}", flush=True)
    pos, tris = scenes.clusters(4000, k=5, seed=27)
    if params["c_node"] > 1e30:
        pos = pos * np.float32(4.0)
    got, ref, _, _ = _collapse_both(ctx, pos, tris, **params)
    _assert_same(str(params), got, ref)


def test_collapse_depth_cut_and_errors(ctx):
    print("[case] wide8 collapse of depth-cut leaves; too many objects per leaf", flush=True)
    pos, tris = scenes.soup(400, seed=17)
    got, ref, (nodes, lens, _, _), _ = _collapse_both(ctx, pos, tris, max_depth=9)
    assert lens[nodes["child_node"] == 0].max() == 3
    _assert_same("depth 9", got, ref)
    t = tris.copy()
    nodes, lens, _, _, _ = ctx.build_tris(pos, t, N.default_settings(max_depth=4))
    d_n, d_l = ctx.to_device(nodes), ctx.to_device(lens)
    d_w, d_t = ctx.alloc(80 * 400), ctx.alloc(4 * 400)
    with pytest.raises(N.MgbvhError) as e:
        ctx.wide8_collapse(d_n, d_l, nodes.shape[0], d_w, 400, d_t, 400)
    assert e.value.code == N.ERR_INVALID_ARG
    with pytest.raises(N.MgbvhError):
        ctx.wide8_collapse(d_n, None, nodes.shape[0], d_w, 400, d_t, 400, p_max=4)
    for bad in (dict(c_node=np.inf), dict(c_prim=np.inf), dict(c_node=np.nan)):
        with pytest.raises(N.MgbvhError) as e:
            ctx.wide8_collapse(d_n, d_l, nodes.shape[0], d_w, 400, d_t, 400, **bad)
        assert e.value.code == N.ERR_INVALID_ARG
    with pytest.raises(N.MgbvhError):  # capacity below the node count
        t = tris.copy()
        full, _, _, _, _ = ctx.build_tris(pos, t, N.default_settings())
        d_f = ctx.to_device(full)
        ctx.wide8_collapse(d_f, None, full.shape[0], d_w, 3, d_t, 400)
    for b in (d_n, d_l, d_w, d_t, d_f):
        b.free()


def _bfs_renumber(nodes, lens):
    """The same tree numbered breadth first (child pairs stay adjacent and in
    order, leaves keep their start indices): children still follow their
    parents, but a subtree's nodes no longer share one index range."""
    old_child = nodes["child_node"].astype(np.int64)
    order, q = [0], 0
    while q < len(order):
        v = order[q]
        q += 1
        c = old_child[v]
        if c:
            order += [c, c + 1]
    order = np.array(order, np.int64)
    new_of = np.empty(len(order), np.int64)
    new_of[order] = np.arange(len(order))
    out = nodes[order].copy()
    ch = old_child[order]
    out["child_node"] = np.where(ch != 0, new_of[np.maximum(ch, 0)], 0).astype(np.uint32)
    return out, lens[order].copy()


def test_collapse_of_other_numberings(ctx):
    # a tree whose subtrees are not contiguous index ranges takes the level
    # path of the collapse; its wide nodes are those of the same topology
    print("[case] wide8 collapse of a breadth-first numbered tree (level path)", flush=True)
    for pos, tris in (scenes.soup(3000, seed=51), scenes.heightfield(50, 40, seed=52)):
        t = tris.copy()
        nodes, lens, _, _, _ = ctx.build_tris(pos, t, N.default_settings())
        bn, bl = _bfs_renumber(nodes, lens)
        assert not np.array_equal(bn["child_node"], nodes["child_node"])
        n = tris.shape[0]
        d_nodes = ctx.to_device(bn)
        d_wide, d_tri = ctx.alloc(N.WIDE_NODE_BYTES * n), ctx.alloc(4 * n)
        nw, cost = ctx.wide8_collapse(d_nodes, None, bn.shape[0], d_wide, n, d_tri, n)
        wide = d_wide.download(np.uint8, N.WIDE_NODE_BYTES * nw).reshape(nw, N.WIDE_NODE_BYTES)
        ref = O.wide8_collapse(bn.view(np.uint32).reshape(-1, 8), None)
        _assert_same("breadth first", (wide, d_tri.download(np.uint32, n), cost), ref)
        # and the same wide nodes as the builder's numbering gives
        rw0, rt0, _ = O.wide8_collapse(nodes.view(np.uint32).reshape(-1, 8), None)
        assert np.array_equal(wide, rw0)
        for b in (d_nodes, d_wide, d_tri):
            b.free()


def test_heightfield_1m_collapse_and_traversal(ctx):
    print("[case] wide8 on the 1024x512 heightfield: collapse vs oracle, 262k camera rays",
          flush=True)
    pos, tris = scenes.heightfield(1024, 512, seed=2)
    got, ref, (nodes, lens, perm, tp), ms = _collapse_both(ctx, pos, tris)
    _assert_same("heightfield 1M", got, ref)
    wide, tri_index, _ = got
    nw = wide.shape[0]
    print(f"  {tris.shape[0]} objects -> {nw} wide nodes, collapse call {ms:.1f} ms", flush=True)
    w, h = 512, 512
    n = w * h
    cam = look_at_camera((0.5, 0.9, -0.6), (0.5, 0.0, 0.5), 0.9, 1.0)
    d_r = ctx.alloc(32 * n)
    ctx.gen_primary(cam, w, h, d_r)
    d_pos, d_tp, d_perm = ctx.to_device(pos), ctx.to_device(tp), ctx.to_device(perm)
    d_nodes, d_wide, d_tri = ctx.to_device(nodes), ctx.to_device(wide), ctx.to_device(tri_index)
    d_h0, d_h1 = ctx.alloc(16 * n), ctx.alloc(16 * n)
    ctx.trace_closest(d_nodes, None, d_pos, 12, d_tp, d_perm, d_r, n, d_h0)
    ctx.trace_wide8(d_wide, nw, d_tri, d_pos, 12, d_tp, d_perm, d_r, n, d_h1)
    a, b = d_h0.download(N.HIT_DTYPE, n), d_h1.download(N.HIT_DTYPE, n)
    assert (a["prim"] != N.MISS).sum() > n // 4
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert np.array_equal(a["prim"], b["prim"]) and np.array_equal(a["tri_slot"], b["tri_slot"])
    for name, run in (("binary", lambda: ctx.trace_closest(d_nodes, None, d_pos, 12, d_tp, d_perm,
                                                           d_r, n, d_h0)),
                      ("wide8", lambda: ctx.trace_wide8(d_wide, nw, d_tri, d_pos, 12, d_tp, d_perm,
                                                        d_r, n, d_h1))):
        ctx.record(0)
        for _ in range(3):
            run()
        ctx.record(1)
        ctx.synchronize()
        print(f"  {name} traversal {ctx.elapsed_ms(0, 1) / 3:.3f} ms for {n} rays", flush=True)
    for buf in (d_r, d_pos, d_tp, d_perm, d_nodes, d_wide, d_tri, d_h0, d_h1):
        buf.free()


def test_wide_traversal_random_rays(ctx):
    print("[case] wide8 traversal vs binary on random and axis-aligned rays", flush=True)
    pos, tris = scenes.clusters(6000, k=9, seed=28)
    got, ref, (nodes, lens, perm, tp), _ = _collapse_both(ctx, pos, tris)
    wide, tri_index, _ = got
    rng = np.random.default_rng(29)
    o = rng.uniform(-0.5, 1.5, (20000, 3))
    d = pos[rng.integers(0, pos.shape[0], 20000)] - o
    d[:500] = [0, 0, 1]
    o[:500, 2] = -1.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = N.make_rays(o.astype(np.float32), d.astype(np.float32))
    d_r = ctx.to_device(rays)
    d_pos, d_tp, d_perm = ctx.to_device(pos), ctx.to_device(tp), ctx.to_device(perm)
    d_nodes, d_wide, d_tri = ctx.to_device(nodes), ctx.to_device(wide), ctx.to_device(tri_index)
    d_h0, d_h1 = ctx.alloc(16 * 20000), ctx.alloc(16 * 20000)
    ctx.trace_closest(d_nodes, None, d_pos, 12, d_tp, d_perm, d_r, 20000, d_h0)
    ctx.trace_wide8(d_wide, wide.shape[0], d_tri, d_pos, 12, d_tp, d_perm, d_r, 20000, d_h1)
    a, b = d_h0.download(N.HIT_DTYPE, 20000), d_h1.download(N.HIT_DTYPE, 20000)
    assert (a["prim"] != N.MISS).sum() > 2000
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert np.array_equal(a["prim"], b["prim"])
    for buf in (d_r, d_pos, d_tp, d_perm, d_nodes, d_wide, d_tri, d_h0, d_h1):
        buf.free()


def test_batch_collapse_equals_per_blas(ctx):
    print("[case] wide8 batch collapse of a 7-mesh batched build (one empty, one 1-triangle)",
          flush=True)
    from marigold_amd.scene import GpuScene
    meshes = [scenes.soup(900, seed=40, extent=0.1), scenes.heightfield(24, 20, seed=41),
              (np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32)),
              scenes.soup(37, seed=42, extent=0.2), scenes.soup(1, seed=44, extent=0.3),
              scenes.clusters(3000, k=5, seed=45), scenes.soup(2, seed=46, extent=0.3)]
    T = np.tile(np.eye(4, dtype=np.float32), (3, 1, 1))
    sc = GpuScene(ctx, meshes, T, np.array([0, 1, 5], np.uint32))
    d = sc.d
    nt = sc.n_tris
    d_wide, d_tri = ctx.alloc(N.WIDE_NODE_BYTES * nt), ctx.alloc(4 * nt)
    first, count, total = ctx.wide8_collapse_batch(d["blas_nodes"], None, sc.n_slots, sc.blas,
                                                   sc.blas_node_counts, d_wide, nt, d_tri, nt)
    wide = d_wide.download(np.uint8, N.WIDE_NODE_BYTES * total).reshape(total, -1)
    tri_index = d_tri.download(np.uint32, nt)
    nodes = sc.download("blas_nodes", np.uint32, 8 * sc.n_slots).reshape(-1, 8)
    assert count[2] == 0 and total == int(count.sum())
    # a batch whose only BLAS is empty: no launch, zero counts
    f2, c2, t2 = ctx.wide8_collapse_batch(d["blas_nodes"], None, sc.n_slots, sc.blas[2:3],
                                          sc.blas_node_counts[2:3], d_wide, nt, d_tri, nt)
    assert t2 == 0 and f2.tolist() == [0] and c2.tolist() == [0]
    for m in range(len(meshes)):
        n = int(sc.counts[m])
        if n == 0:
            continue
        off = int(sc.node_offsets[m])
        rw, rt, _ = O.wide8_collapse(nodes[off:off + 2 * n - 1])
        got = wide[first[m]:first[m] + count[m]]
        assert got.shape == rw.shape and np.array_equal(got, rw), m
        to = int(sc.tri_offsets[m])
        assert np.array_equal(tri_index[to:to + n], rt), m
    d_wide.free()
    d_tri.free()
    sc.free()
