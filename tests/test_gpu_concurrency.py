"""Concurrent builds from several host threads, one context each.

marigold builds its BLASes from rayon workers at once (model.rs:77-80), which
is why BoundingVolumeHierarchy::new requires S: Sync, T: Sync (lib.rs:698).
Here 6 threads, each with its own context (own stream and scratch), call the C
ABI's mgbvh_build_tris on different meshes at the same time, three times each;
every result must equal the oracle's tree, permutation and split records.
"""
import threading

import numpy as np
import pytest

from marigold_amd import _native as N
from marigold_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_concurrent_contexts_equal_oracle():
    print("[case] 6 threads, one context each, 3 builds per thread", flush=True)
    meshes = [scenes.soup(6000 + 4000 * i, seed=70 + i) for i in range(3)]
    meshes += [scenes.heightfield(60 + 20 * i, 40, seed=80 + i) for i in range(3)]
    refs = []
    for pos, tris in meshes:
        r = O.build(O.tri_aabbs(pos, tris), max_depth=32)
        refs.append((O.encode(r["nodes"]), r["list"], r["recs"].view(np.uint32).reshape(-1, 4)))
    start = threading.Barrier(len(meshes))
    errors = [None] * len(meshes)

    def work(i):
        try:
            pos, tris = meshes[i]
            want_nodes, want_perm, want_recs = refs[i]
            with N.Context(0) as c:
                start.wait(timeout=60)
                for _ in range(3):
                    t = tris.copy()
                    nodes, lens, perm, splits, _ = c.build_tris(pos, t, N.default_settings(),
                                                                want_splits=True)
                    assert np.array_equal(nodes.view(np.uint32).reshape(-1, 8), want_nodes)
                    assert np.array_equal(perm, want_perm)
                    assert np.array_equal(t, tris[want_perm])
                    assert np.array_equal(splits.view(np.uint32).reshape(-1, 4), want_recs)
        except BaseException as e:  # reported by the main thread
            errors[i] = e

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(meshes))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
        assert not th.is_alive(), "a build thread did not finish"
    for i, e in enumerate(errors):
        assert e is None, (i, repr(e))
