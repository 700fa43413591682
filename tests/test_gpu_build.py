"""GPU parity of the plain builder against the oracle, through the C ABI.

Bit-exact: 32-byte node records (topology, start indices, fp32 bounds), the
object permutation, and the per-node split records (axis, method, bin cut or
threshold bits, count_lt, cost bits).
"""
import numpy as np
import pytest

from marigold_amd import _native as N
from marigold_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _check_tris(ctx, pos, tris, max_depth=32, bounds=None, tag=""):
    print(f"[case] {tag} n={tris.shape[0]} max_depth={max_depth} bounds={bounds is not None}",
          flush=True)
    ref = O.build(O.tri_aabbs(pos, tris), max_depth=max_depth, bounds=bounds)
    t = tris.copy()
    nodes, lens, perm, splits, stats = ctx.build_tris(
        pos, t, N.default_settings(max_depth=max_depth, bounds=bounds), want_splits=True)
    want = O.encode(ref["nodes"])
    got = nodes.view(np.uint32).reshape(-1, 8)
    assert got.shape == want.shape, (tag, got.shape, want.shape)
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert bad.size == 0, (tag, "first differing node", bad[:5], got[bad[:3]], want[bad[:3]])
    assert np.array_equal(lens, ref["nodes"]["len"]), tag
    assert np.array_equal(perm, ref["list"]), tag
    assert np.array_equal(t, tris[ref["list"]]), tag
    rs = ref["recs"].view(np.uint32).reshape(-1, 4)
    gs = splits.view(np.uint32).reshape(-1, 4)
    bad = np.nonzero((gs != rs).any(axis=1))[0]
    assert bad.size == 0, (tag, "first differing split record", bad[:5], gs[bad[:3]], rs[bad[:3]])
    assert stats["height"] == ref["stats"]["height"], tag
    assert stats["node_count"] == len(ref["nodes"]), tag
    return ref, stats


def test_centroids_on_bin_boundaries(ctx):
    # point triangles whose x centroids sit on k/32 of the root's centroid range
    # and one ulp either side: the multiply-by-reciprocal bin index must hand
    # these to the exact division (root: several chunks; children: one chunk)
    rng = np.random.default_rng(21)
    base = np.arange(33, dtype=np.float32) / np.float32(32)
    xs = np.concatenate([base, np.nextafter(base, np.float32(-1)), np.nextafter(base, np.float32(2))])
    xs = np.clip(xs, np.float32(0), np.float32(1)).astype(np.float32)
    n = 5000
    x = xs[rng.integers(0, xs.size, n)]
    x[:2] = [0.0, 1.0]
    pts = np.stack([x, rng.random(n, dtype=np.float32), rng.random(n, dtype=np.float32)], axis=1)
    pos = np.repeat(pts, 3, axis=0).astype(np.float32)
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    _check_tris(ctx, pos, tris, tag="bin-boundaries")


def test_soup_65536_bit_exact(ctx):
    pos, tris = scenes.soup(65536, seed=1)
    ref, st = _check_tris(ctx, pos, tris, tag="soup65536")
    assert ref["stats"]["n_tie_events"] == 0


def test_heightfield_1m_bit_exact(ctx):
    pos, tris = scenes.heightfield(1024, 512, seed=2)
    ref, st = _check_tris(ctx, pos, tris, tag="heightfield1024x512")
    assert ref["stats"]["n_tie_events"] == 0
    assert ref["stats"]["n_median"] > 0  # coincident cell pairs exercise the median fallback


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 9, 16, 31, 32, 33, 34, 63, 64, 65, 100, 257,
                               2047, 2048, 2049, 4097, 10000])
def test_soup_sizes(ctx, n):
    pos, tris = scenes.soup(n, seed=100 + n)
    _check_tris(ctx, pos, tris, tag=f"soup{n}")


@pytest.mark.parametrize("max_depth", [0, 1, 2, 5, 9, 12])
def test_depth_cutoff(ctx, max_depth):
    pos, tris = scenes.soup(3000, seed=5)
    ref, st = _check_tris(ctx, pos, tris, max_depth=max_depth, tag=f"cutoff{max_depth}")
    if max_depth < 10:
        assert ref["stats"]["n_cutoff_leaves"] > 0


def test_given_bounds(ctx):
    pos, tris = scenes.soup(5000, seed=9)
    lo = pos.min(axis=0) - np.float32(0.25)
    hi = pos.max(axis=0) + np.float32(0.5)
    _check_tris(ctx, pos, tris, bounds=(lo, hi), tag="bounds")


def test_identical_triangles_median(ctx):
    # every centroid coincides: binned and adaptive sweeps refuse, median fallback
    for n in (2, 5, 31, 40, 3000):
        pos = np.tile(np.array([[0.1, 0.2, 0.3], [0.4, 0.2, 0.3], [0.1, 0.6, 0.9]], np.float32), (n, 1))
        tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
        ref, _ = _check_tris(ctx, pos, tris, tag=f"identical{n}")
        assert ref["stats"]["n_median"] == n - 1


def test_collinear_and_coplanar(ctx):
    n = 500
    x = np.linspace(0.0, 1.0, n, dtype=np.float32)
    pos = np.zeros((3 * n, 3), np.float32)
    pos[0::3, 0] = x
    pos[1::3, 0] = x + np.float32(0.001)
    pos[2::3, 0] = x
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    _check_tris(ctx, pos, tris, tag="collinear")
    pos2 = pos.copy()
    pos2[2::3, 1] = np.float32(0.002)
    _check_tris(ctx, pos2, tris, tag="coplanar")


def test_tiny_extents(ctx):
    # centroids a few ULPs apart: adaptive thresholds can all round onto one side
    base = np.float32(1.0)
    ulp = np.spacing(base)
    for n in (9, 20, 32, 33, 80):
        rng = np.random.default_rng(n)
        off = rng.integers(0, 3, size=(n, 3)).astype(np.float32) * ulp
        c = base + off
        pos = np.repeat(c, 3, axis=0).astype(np.float32)
        tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
        _check_tris(ctx, pos, tris, tag=f"tiny{n}")


def test_aabb_objects_tlas(ctx):
    rng = np.random.default_rng(3)
    m = 10000
    c = rng.uniform(-100, 100, size=(m, 3)).astype(np.float32)
    h = rng.uniform(0.5, 3.0, size=(m, 3)).astype(np.float32)
    boxes = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    print("[case] tlas10000", flush=True)
    ref = O.build(boxes, max_depth=32)
    nodes, lens, perm, splits, stats = ctx.build_aabbs(boxes, N.default_settings(), want_splits=True)
    assert np.array_equal(nodes.view(np.uint32).reshape(-1, 8), O.encode(ref["nodes"]))
    assert np.array_equal(perm, ref["list"])
    assert np.array_equal(splits.view(np.uint32).reshape(-1, 4), ref["recs"].view(np.uint32).reshape(-1, 4))


def test_input_order_invariance(ctx):
    # the node array is a function of the geometry, not of the caller's order
    pos, tris = scenes.heightfield(64, 32, seed=11)
    s = N.default_settings()
    a = ctx.build_tris(pos, tris.copy(), s)[0]
    rng = np.random.default_rng(0)
    sh = tris[rng.permutation(tris.shape[0])].copy()
    b = ctx.build_tris(pos, sh, s)[0]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_empty_and_errors(ctx):
    pos = np.zeros((3, 3), np.float32)
    tris = np.zeros((0, 3), np.uint32)
    nodes, lens, perm, splits, stats = ctx.build_tris(pos, tris, N.default_settings())
    assert len(nodes) == 0
    bad = np.array([[0, 1, 7]], np.uint32)
    with pytest.raises(N.MgbvhError) as e:
        ctx.build_tris(pos, bad.copy(), N.default_settings())
    assert e.value.code == N.ERR_INDEX
    pos2 = np.array([[0, 0, 0], [1, np.nan, 0], [0, 1, 0]], np.float32)
    with pytest.raises(N.MgbvhError) as e:
        ctx.build_tris(pos2, np.array([[0, 1, 2]], np.uint32), N.default_settings())
    assert e.value.code == N.ERR_NONFINITE
    s = N.default_settings()
    s.max_leaf_objects = 4
    with pytest.raises(N.MgbvhError) as e:
        ctx.build_tris(pos, np.array([[0, 1, 2]], np.uint32), s)
    assert e.value.code == N.ERR_UNSUPPORTED


def test_repeatable_device_entry(ctx):
    """Device-resident entry point gives the same tree on repeated calls."""
    pos, tris = scenes.soup(20000, seed=21)
    n = tris.shape[0]
    d_pos, d_tris = ctx.to_device(pos), ctx.to_device(tris)
    d_nodes = ctx.alloc((2 * n - 1) * 32)
    d_perm = ctx.alloc(n * 4)
    d_tout = ctx.alloc(n * 12)
    import ctypes as C
    outs = []
    for _ in range(3):
        nn = C.c_uint64()
        N.check(N.lib().mgbvh_build_tris_device(ctx.handle, d_pos.ptr, 12, pos.shape[0], d_tris.ptr, n,
                                                C.byref(N.default_settings()), d_nodes.ptr, None,
                                                d_perm.ptr, d_tout.ptr, None, C.byref(nn), None))
        outs.append((d_nodes.download(np.uint32, (2 * n - 1) * 8), d_perm.download(np.uint32, n)))
    ref = O.build(O.tri_aabbs(pos, tris))
    for nd, pm in outs:
        assert np.array_equal(nd.reshape(-1, 8), O.encode(ref["nodes"]))
        assert np.array_equal(pm, ref["list"])
    assert np.array_equal(d_tout.download(np.uint32, 3 * n).reshape(-1, 3), tris[ref["list"]])
