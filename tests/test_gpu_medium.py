"""The plain builder with nodes of 33..medium_max_objects built one wavefront
per subtree (k_medium) instead of level by level, and with that off (1, every
node > 32 objects on the level path): the same bit-exact checks against the
oracle as tests/test_gpu_build.py (nodes, lengths, permutation, permuted
triangles, split records), run with each tuning."""
import numpy as np
import pytest

import test_gpu_build as B
from marigold_amd import _native as N
from marigold_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    "bin-boundaries": lambda ctx: B.test_centroids_on_bin_boundaries(ctx),
    "soup65536": lambda ctx: B.test_soup_65536_bit_exact(ctx),
    "heightfield1m": lambda ctx: B.test_heightfield_1m_bit_exact(ctx),
    "soup-sizes": lambda ctx: [B.test_soup_sizes(ctx, n) for n in (33, 34, 64, 65, 511, 512, 513,
                                                                  2047, 2048, 2049, 10000)],
    "depth-cut": lambda ctx: [B.test_depth_cutoff(ctx, d) for d in (1, 2, 5, 9, 14)],
    "bounds": lambda ctx: B.test_given_bounds(ctx),
    "identical": lambda ctx: B.test_identical_triangles_median(ctx),
    "collinear": lambda ctx: B.test_collinear_and_coplanar(ctx),
    "tiny": lambda ctx: B.test_tiny_extents(ctx),
    "aabbs": lambda ctx: B.test_aabb_objects_tlas(ctx),
    "order": lambda ctx: B.test_input_order_invariance(ctx),
}


@pytest.fixture(params=[1, 512, 2048])
def medium(ctx, request):
    old = ctx.get_tuning()
    ctx.set_tuning(medium_max_objects=request.param)
    assert ctx.get_tuning()["medium_max_objects"] == request.param  # 1: off, level path only
    yield request.param
    ctx.set_tuning(**{k: v for k, v in old.items() if k != "compute_units"})


@pytest.mark.parametrize("case", list(CASES))
def test_medium_builds_bit_exact(ctx, medium, case):
    print(f"[case] medium_max_objects={medium} {case}", flush=True)
    CASES[case](ctx)


def test_medium_stats_and_clusters(ctx, medium):
    # 40,000 clustered triangles: stats count the medium splits as binned nodes
    print(f"[case] medium_max_objects={medium} clusters 40000 stats", flush=True)
    pos, tris = scenes.clusters(40000, k=9, seed=61)
    ref, st = B._check_tris(ctx, pos, tris, tag=f"clusters-medium{medium}")
    lens = ref["nodes"]["len"]
    internal = ref["nodes"]["child"] != 0
    big = internal & (lens > 32)
    assert st["large_nodes"] == int(big.sum())
    assert st["large_object_visits"] == int(lens[big].astype(np.int64).sum())


def test_medium_tuning_bounds(ctx):
    with pytest.raises(N.MgbvhError):
        ctx.set_tuning(medium_max_objects=32)
    with pytest.raises(N.MgbvhError):
        ctx.set_tuning(medium_max_objects=2049)
