"""The sample collector (marigold_amd/sample_collector.py, the command line of
bvh_sample_collector/src/main.rs:9-70) on the GPU: glTF scenes in, one
profiling JSON per build out (bvh/src/lib.rs:836-866), every tree the oracle's
for the mesh's own bounds and max_depth 24."""
import glob
import json
import os

import numpy as np
import pytest

from marigold_amd import _native as N
from marigold_amd import gltf, scenes
from marigold_amd.sample_collector import collect, main
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _records(pos, tris):
    v = np.zeros((pos.shape[0], 8), np.float32)
    v[:, 0:3] = pos
    v[:, 5] = 1.0
    return v, tris


def test_collector_writes_one_json_per_build(ctx, tmp_path):
    print("[case] sample collector: 2 glTF scenes, 3 meshes, 2 samples", flush=True)
    p0, t0 = scenes.soup(3000, seed=71)
    # an unreferenced far vertex: the accessor bounds exceed the triangles' union
    p0 = np.concatenate([p0, np.float32([[4.0, -3.0, 2.0]])])
    p1, t1 = scenes.heightfield(40, 25, seed=72)
    p2, t2 = scenes.clusters(700, k=3, seed=73)
    mesh_dir = tmp_path / "meshes"
    gltf.write(str(mesh_dir / "alpha"), [_records(p0, t0), _records(p1, t1)], [("s", [])])
    gltf.write(str(mesh_dir / "beta"), [_records(p2, t2)], [("s", [])])
    out = tmp_path / "snapshots"
    lines = []
    assert collect(str(mesh_dir), str(out), "gen1", 2, ctx=ctx, out=lines.append) == 6
    assert lines[0] == "loading gltf scene alpha..." and lines[-1] == "\tsample 2 done for beta!"
    files = sorted(glob.glob(str(out / "gen1" / "bvh_*.json")))
    assert len(files) == 6
    recs = [json.loads(open(f).read()) for f in files]
    by_name = {}
    for r in recs:
        by_name.setdefault(r["name"], []).append(r)
    assert sorted(by_name) == ["alpha_0", "alpha_1", "beta_0"]
    for name, (pos, tris) in (("alpha_0", (p0, t0)), ("alpha_1", (p1, t1)), ("beta_0", (p2, t2))):
        n = tris.shape[0]
        for r in by_name[name]:
            assert r["node_count"] == 2 * n - 1 and r["total_object_count"] == n
            assert r["max_depth"] == 24 and r["height"] <= 24
            assert r["min_leaf_object_count"] >= 1
        # the collector's settings give the oracle's tree for these bounds
        lo, hi = pos.min(0), pos.max(0)
        t = tris.copy()
        nodes, lens, perm, _, _ = ctx.build_tris(
            _records(pos, tris)[0], t, N.default_settings(max_depth=24, bounds=(lo, hi)))
        ref = O.build(O.tri_aabbs(pos, tris), max_depth=24, bounds=(lo, hi))
        assert np.array_equal(nodes.view(np.uint32).reshape(-1, 8), O.encode(ref["nodes"]))
        assert np.array_equal(perm, ref["list"])


def test_collector_usage(capsys):
    assert main(["only", "two"]) == 2
    assert "sample_collector" in capsys.readouterr().err
