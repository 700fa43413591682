"""Pins of two oracle functions that the other oracle tests leave open.

1. The adaptive sweep's uniform-threshold branch for 8 <= n <= 32 objects
   (bvh/src/lib.rs:428-441): S = n thresholds t_i = cmin + (cext / S) * (i + 0.5).
   The check enumerates those thresholds independently in numpy float32 (the
   reference's operation order), forms every partition {c < t_i} on all three
   axes and prices it with exact rational SAH, then requires the oracle's root
   decision (axis, threshold bits, count_lt) to be one of those candidates, of
   minimum cost, and the first such in (axis, i) order when the minimum is not
   a near tie.
2. The tree SAH cost (or_tree_sah) with the reference's constants
   DEPTH_COST = 1 and OBJECT_COST = 8 (lib.rs:217-218):
   C = sum_internal SA(v)/SA(root) + sum_leaf 8 len(l) SA(l)/SA(root),
   against hand-derived known answers.

Each check is shown to fail on planted mistakes built into separate variants
of the oracle (OR_PLANT in oracle/bvh_oracle.c): S = 32 for every n, the
+ 0.5 dropped, the leaf term without len, and the internal term weighted 8.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

ADAPTIVE_SIZES = (8, 9, 16, 31, 32)
SEEDS = range(10)


def _sa(b):
    w, h, d = (Fraction(float(b[3 + a])) - Fraction(float(b[a])) for a in range(3))
    return 2 * (w * h + w * d + h * d) if min(w, h, d) >= 0 else Fraction(0)


def _sah(parent, lt, gt, nl, ng):
    return 1 + _sa(lt) / _sa(parent) * 8 * nl + _sa(gt) / _sa(parent) * 8 * ng


def _union(b):
    return np.concatenate([b[:, :3].min(axis=0), b[:, 3:].max(axis=0)])


def _boxes(n, seed):
    rng = np.random.default_rng(7919 * n + seed)
    c = rng.random((n, 3)).astype(np.float32)
    e = (rng.random((n, 3)) * 0.2 + 0.001).astype(np.float32)
    return np.concatenate([c - e, c + e], axis=1).astype(np.float32)


def _thresholds(cen_a, n):
    """t_i of lib.rs:430-437 in float32, each op rounded as the reference does."""
    f = np.float32
    cmin, cmax = cen_a.min(), cen_a.max()
    step = f(f(cmax - cmin) / f(n))
    return [f(cmin + f(step * f(f(i) + f(0.5)))) for i in range(n)]


def _check_adaptive_root(n, seed, L=None):
    """Raises AssertionError when the root decision is not the first minimum
    over the uniform thresholds.  Returns True when the first-minimum rule was
    decided (no near tie between different partitions)."""
    boxes = _boxes(n, seed)
    r = O.build(boxes, L=L)
    rec = r["recs"][0]
    cen = ((boxes[:, :3] + boxes[:, 3:]) * np.float32(0.5)).astype(np.float32)
    parent = _union(boxes)
    cands = []  # (exact cost, axis, i, threshold bits, lt mask)
    for a in range(3):
        for i, t in enumerate(_thresholds(cen[:, a], n)):
            m = cen[:, a] < t
            k = int(m.sum())
            if 0 < k < n:  # MIN_LEAF_OBJECTS = 1 on both sides (lib.rs:407-409)
                cost = _sah(parent, _union(boxes[m]), _union(boxes[~m]), k, n - k)
                cands.append((cost, a, i, int(np.array([t], np.float32).view(np.uint32)[0]), m))
    assert cands
    axis, kind = rec["axis_method"] & 0xFF, rec["axis_method"] >> 8
    assert kind == 1, "an 8..32-object node must be split by the adaptive sweep"
    mine = [cd for cd in cands if cd[1] == axis and cd[3] == int(rec["value"])]
    assert mine, "the recorded threshold is not one of the sweep's thresholds on its axis"
    cost, _, _, _, m = mine[0]
    assert int(rec["count_lt"]) == int(m.sum())
    best = min(cd[0] for cd in cands)
    assert cost <= best * (1 + Fraction(1, 10 ** 5)), "the recorded split is not of minimum SAH"
    # first minimum in (axis, i) order, decided when every candidate within the
    # float32 rounding margin of the best has the same partition
    near = [cd for cd in cands if cd[0] <= best * (1 + Fraction(1, 10 ** 5))]
    if all(np.array_equal(cd[4], near[0][4]) and cd[1] == near[0][1] for cd in near):
        first = near[0]
        assert (axis, int(rec["value"])) == (first[1], first[3]), "not the first minimum"
        return True
    return False


@pytest.mark.parametrize("n", ADAPTIVE_SIZES)
def test_adaptive_uniform_branch_brute_force(n):
    decided = sum(_check_adaptive_root(n, s) for s in SEEDS)
    assert decided >= len(SEEDS) - 2  # near ties are rare on random boxes


@pytest.mark.parametrize("plant", [1, 2])
def test_adaptive_check_fails_on_planted_mistakes(plant):
    """S = 32 for every n (plant 1) and thresholds without + 0.5 (plant 2)
    must each fail the brute-force check on some size."""
    L = O.planted_lib(plant)
    fails = 0
    for n in ADAPTIVE_SIZES:
        for s in SEEDS:
            try:
                _check_adaptive_root(n, s, L=L)
            except AssertionError:
                fails += 1
    assert fails >= 10, f"planted variant {plant} failed only {fails} cases"


def test_adaptive_sweep_only_for_32_or_fewer():
    """33 objects take the binned sweep (kind 0), 32 the adaptive one (kind 1)."""
    assert O.build(_boxes(33, 0))["recs"][0]["axis_method"] >> 8 == 0
    assert O.build(_boxes(32, 0))["recs"][0]["axis_method"] >> 8 == 1


# ---------------------------------------------------------------- tree SAH
def _nodes(rows):
    """rows: (min3, max3, start, len, child) -> oracle node array."""
    a = np.zeros(len(rows), dtype=O.NODE_DTYPE)
    for k, (mn, mx, st, ln, ch) in enumerate(rows):
        a[k] = (mn, mx, st, ln, ch)
    return a


def _cube(x):
    return ([x, 0, 0], [x + 1, 1, 1])


# A = [0,1]^3, B = [2,3]x[0,1]^2: root [0,3]x[0,1]^2 (SA 14), leaves SA 6:
# 14/14 + 8 (6 + 6) / 14 = 55/7.
TWO = _nodes([([0, 0, 0], [3, 1, 1], 0, 2, 1), (*_cube(2), 1, 1, 0), (*_cube(0), 0, 1, 0)])
TWO_SAH = Fraction(55, 7)
# A, B, C at x = 0, 2, 5: root [0,6] (SA 26) splits {A,B} | {C}; {A,B} = [0,3]
# (SA 14) splits into B (gt) and A (lt): (26 + 14)/26 + 8 (6 + 6 + 6)/26 = 92/13.
THREE = _nodes([([0, 0, 0], [6, 1, 1], 0, 3, 1), (*_cube(5), 2, 1, 0),
                ([0, 0, 0], [3, 1, 1], 0, 2, 3), (*_cube(2), 1, 1, 0), (*_cube(0), 0, 1, 0)])
THREE_SAH = Fraction(92, 13)
# the same objects cut at depth 1: {A,B} becomes a leaf of 2 objects:
# 26/26 + 8 (6 + 2 * 14)/26 = 149/13.
CUT = _nodes([([0, 0, 0], [6, 1, 1], 0, 3, 1), (*_cube(5), 2, 1, 0),
              ([0, 0, 0], [3, 1, 1], 0, 2, 0)])
CUT_SAH = Fraction(149, 13)
KATS = [(TWO, TWO_SAH), (THREE, THREE_SAH), (CUT, CUT_SAH)]


def _close(x, q):
    return abs(Fraction(x) - q) <= q * Fraction(1, 10 ** 12)


@pytest.mark.parametrize("k", range(len(KATS)))
def test_tree_sah_known_answers(k):
    nodes, want = KATS[k]
    assert _close(O.tree_sah(nodes), want)


def test_tree_sah_single_leaf_and_flat_root():
    one = _nodes([(*_cube(0), 0, 1, 0)])
    assert O.tree_sah(one) == 8.0        # 8 * 1 * SA/SA
    flat = _nodes([([0, 0, 0], [0, 0, 0], 0, 1, 0)])
    assert O.tree_sah(flat) == 0.0       # SA(root) = 0: no normaliser


def test_builder_makes_the_hand_derived_trees():
    """The oracle's own builds of the three-cube scene are the trees above
    (adaptive n < 8 sweep: {A,B}|{C} costs 1 + 272/26 against 1 + 336/26)."""
    boxes = np.array([[0, 0, 0, 1, 1, 1], [2, 0, 0, 3, 1, 1], [5, 0, 0, 6, 1, 1]], np.float32)
    for depth, want in ((32, THREE), (1, CUT)):
        r = O.build(boxes, max_depth=depth)
        got = r["nodes"]
        for f in ("min", "max", "start", "len", "child"):
            assert np.array_equal(got[f], want[f]), (depth, f)
        assert list(r["list"]) == [0, 1, 2]
    r = O.build(boxes[:2])
    for f in ("min", "max", "len", "child"):
        assert np.array_equal(r["nodes"][f], TWO[f]), f


@pytest.mark.parametrize("plant", [3, 4])
def test_tree_sah_check_fails_on_planted_mistakes(plant):
    """The leaf term without len (plant 3) or an internal weight of 8 (plant 4)
    changes at least one known answer."""
    L = O.planted_lib(plant)
    assert any(not _close(O.tree_sah(nodes, L=L), want) for nodes, want in KATS)
