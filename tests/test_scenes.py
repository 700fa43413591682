"""The synthetic workloads are deterministic and clean (no NaN, no -0.0)."""
import numpy as np

from marigold_amd import scenes
from oracle import oracle as O


def _clean(pos):
    assert np.isfinite(pos).all()
    assert not np.any(np.signbit(pos) & (pos == 0))


def test_soup_deterministic_and_shaped():
    p1, t1 = scenes.soup(5000, seed=1)
    p2, t2 = scenes.soup(5000, seed=1)
    assert p1.dtype == np.float32 and t1.dtype == np.uint32
    assert p1.shape == (15000, 3) and t1.shape == (5000, 3)
    assert np.array_equal(p1, p2) and np.array_equal(t1, t2)
    _clean(p1)
    c = p1.reshape(-1, 3, 3).mean(axis=1)
    assert c.min() >= -0.02 and c.max() <= 1.02
    off = p1.reshape(-1, 3, 3) - p1.reshape(-1, 3, 3).mean(axis=1, keepdims=True)
    assert np.abs(off).max() <= 0.0201


def test_heightfield_counts_and_range():
    p, t = scenes.heightfield(64, 32, seed=2)
    assert p.shape == (65 * 33, 3) and t.shape == (2 * 64 * 32, 3)
    assert t.max() < p.shape[0]
    _clean(p)
    assert np.abs(p[:, 1]).max() <= 0.2 + 1e-6
    assert p[:, 0].min() == 0.0 and p[:, 0].max() == 1.0
    p2, t2 = scenes.heightfield(64, 32, seed=2)
    assert np.array_equal(p, p2)


def test_instance_boxes_known_answers():
    # translation only: the box moves; a quarter turn about y swaps x and z
    lo = np.array([[0, 0, 0], [1, 2, 3]], np.float32)
    hi = np.array([[1, 2, 3], [2, 4, 6]], np.float32)
    T = np.zeros((2, 3, 4), np.float32)
    T[0] = [[1, 0, 0, 5], [0, 1, 0, -1], [0, 0, 1, 0.5]]
    T[1] = [[0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]]  # x' = z, z' = -x
    b = O.instance_boxes(lo, hi, T.reshape(2, 12))
    assert np.array_equal(b[0], np.float32([5, -1, 0.5, 6, 1, 3.5]))
    assert np.array_equal(b[1], np.float32([3, 2, -2, 6, 4, -1]))


def test_instance_boxes_contain_transformed_corners():
    rng = np.random.default_rng(0)
    lo = rng.random((50, 3)).astype(np.float32)
    hi = lo + rng.random((50, 3)).astype(np.float32)
    T = scenes.random_transforms(rng, 50)
    b = O.instance_boxes(lo, hi, T[:, :3, :].reshape(50, 12))
    for k in range(8):
        c = np.where([(k >> 2) & 1, (k >> 1) & 1, k & 1], hi, lo).astype(np.float64)
        w = np.einsum("mij,mj->mi", T[:, :3, :3].astype(np.float64), c) + T[:, :3, 3]
        assert np.all(w >= b[:, :3] - 1e-4) and np.all(w <= b[:, 3:] + 1e-4)
        # every box face is touched by some corner
    assert np.all(b[:, :3] <= b[:, 3:])