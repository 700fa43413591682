"""CPU pins of the oracle's restatement of the path tracer's own traversal
(scene.slang:91-221, rayIntersectBoundingVolume :296-310, Hit.merge
intersect.slang:73-82), which is tree dependent by design.

Known answers on hand-built trees fix the traversal order, the tie rule (the
hit tested later wins), the object-space distance scaling of :165-166 with no
conversion back (:211-213), the root miss and the shared 64-entry stack.  On
trees the oracle builds, with translation-only instances and unit directions
(no scaling), every hit must be a real hit of the reported triangle at exactly
the reported distance, never closer than the brute-force closest hit, and equal
to it on nearly every ray (the shader's box test is not conservative)."""
import numpy as np
import pytest

from marigold_amd import scenes
from oracle import oracle as O

NONE = 0xFFFFFFFF
TRI = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [0.0, 1.0, 0.0]], np.float32)


def node(bmin, start, bmax, child):
    w = np.zeros(8, np.uint32)
    w[0:3] = np.asarray(bmin, np.float32).view(np.uint32)
    w[3] = start
    w[4:7] = np.asarray(bmax, np.float32).view(np.uint32)
    w[7] = child
    return w


def inst(rows, blas=0):
    r = np.zeros(16, np.uint32)
    r[:12] = np.asarray(rows, np.float32).reshape(12).view(np.uint32)
    r[12] = blas
    return r


def rows(scale, tz):
    return [[scale, 0, 0, 0], [0, scale, 0, 0], [0, 0, scale, tz]]


def ray(o, d):
    return np.array([list(o) + list(d)], np.float32)


def leaf_blas():
    """One triangle in the z = 0 plane; the root is its leaf."""
    return np.stack([node([-1, -1, 0], 0, [1, 1, 0], 0)])


def test_equal_distance_goes_to_the_later_hit():
    # two coincident triangles in sibling leaves with equal boxes: the second
    # child is pushed last, so it is tested first, and the first child's equal
    # hit replaces it (merge keeps a only when strictly closer)
    blas8 = np.stack([node([-1, -1, 0], 0, [1, 1, 0], 1), node([-1, -1, 0], 0, [1, 1, 0], 0),
                      node([-1, -1, 0], 1, [1, 1, 0], 0)])
    tlas8 = np.stack([node([-1, -1, 5], 0, [1, 1, 5], 0)])
    tris = np.array([[0, 1, 2], [0, 1, 2]], np.uint32)
    blas_perm = np.array([1, 0], np.uint32)
    out, rc = O.raytrace_tlas(tlas8, [0], np.stack([inst(rows(1, -5))]), [[0, 0, 0, 0]], blas8, TRI,
                              tris, blas_perm, ray((0, 0, 0), (0, 0, 1)))
    assert rc == 0
    assert out[0].tolist() == [np.float32(5.0).view(np.uint32), 1, 0, 0]


def _one_instance(blas8, pos, blas_perm):
    tlas8 = np.stack([node(blas8[0, 0:3].view(np.float32), 0, blas8[0, 4:7].view(np.float32), 0)])
    tris = np.tile(np.array([[0, 1, 2]], np.uint32), (len(blas_perm), 1))
    out, rc = O.raytrace_tlas(tlas8, [0], np.stack([inst(rows(1, 0))]), [[0, 0, 0, 0]], blas8,
                              pos, tris, np.array(blas_perm, np.uint32), ray((0, 0, 0), (0, 0, 1)))
    assert rc == 0
    return out[0].tolist()


def test_children_at_the_closest_distance_are_culled():
    # leaf 1 (box near 4.5) is popped first and hits at 5; the internal node 2
    # (box near 5) is popped next: its child 3 has box distance 5, not below
    # the closest 5, so its coincident triangle is never tested
    p5 = TRI + np.float32([0, 0, 5])
    blas8 = np.stack([node([-1, -1, 4.5], 0, [1, 1, 6], 1), node([-1, -1, 4.5], 0, [1, 1, 5], 0),
                      node([-1, -1, 5], 0, [1, 1, 6], 3), node([-1, -1, 5], 1, [1, 1, 5], 0),
                      node([-1, -1, 6], 1, [1, 1, 6], 0)])
    assert _one_instance(blas8, p5, [0, 1]) == [np.float32(5).view(np.uint32), 0, 0, 0]


def test_origin_on_a_box_face_takes_the_far_distance():
    # child 1's box starts at the origin's plane (t_near == 0), so its distance
    # is t_far = 3 and child 2 (distance 1) is popped first; child 1's equal
    # hit at t = 2 is tested later and wins the tie
    p2 = TRI + np.float32([0, 0, 2])
    blas8 = np.stack([node([-1, -1, 0], 0, [1, 1, 3], 1), node([-1, -1, 0], 0, [1, 1, 3], 0),
                      node([-1, -1, 1], 1, [1, 1, 2], 0)])
    assert _one_instance(blas8, p2, [0, 1]) == [np.float32(2).view(np.uint32), 0, 0, 0]


def _two_instances(q_tz):
    # P: unit scale, triangle at world z = 5.  Q: half scale (world_to_object
    # 0.5), triangle at world z = q_tz.  P's TLAS box is nearer (3 < 4), so it
    # is popped first; Q's box is still popped (no re-cull after a push).
    tlas8 = np.stack([node([-2, -2, 3], 0, [2, 2, 8], 1), node([-1, -1, 3], 0, [1, 1, 5], 0),
                      node([-2, -2, 4], 1, [2, 2, 8], 0)])
    inst16 = np.stack([inst(rows(1, -5)), inst(rows(0.5, -0.5 * q_tz))])
    return tlas8, inst16


@pytest.mark.parametrize("q_tz,note", [(8.0, "behind"), (4.5, "in front, lost")])
def test_object_space_distance_scaling(q_tz, note):
    tlas8, inst16 = _two_instances(q_tz)
    args = (tlas8, [0, 1], inst16, [[0, 0, 0, 0]], leaf_blas(), TRI, [[0, 1, 2]], [0],
            ray((0, 0, 0), (0, 0, 1)))
    out, rc = O.raytrace_tlas(*args)
    assert rc == 0
    # Q's pass starts from 5 * length(0.5 * d) = 2.5; its own hit (t = 8 or
    # 4.5 as the same ray parameter) is not below 2.5, so the pass returns
    # P's hit with distance 2.5, which then beats P's 5 in the world merge.
    assert out[0].tolist() == [np.float32(2.5).view(np.uint32), 0, 0, 0], note
    # the tree-independent closest hit: P at 5 behind, Q at 4.5 in front
    t, prim, ins = O.closest_hit_scene_brute(TRI, np.array([[0, 1, 2]], np.uint32),
                                             np.array([[0, 1, 0]], np.uint32), inst16,
                                             args[-1])
    assert (float(t[0]), int(ins[0])) == ((5.0, 0) if q_tz > 5 else (4.5, 1))


def test_root_miss_and_unit_scale():
    tlas8, inst16 = _two_instances(8.0)
    base = (tlas8, [0, 1], inst16, [[0, 0, 0, 0]], leaf_blas(), TRI, [[0, 1, 2]], [0])
    out, rc = O.raytrace_tlas(*base, ray((0, 0, 0), (0, 0, -1)))
    assert rc == 0 and out[0].tolist() == [0, NONE, NONE, NONE]  # NO_HIT, distance 0
    # both instances at unit scale: the nearer triangle wins at its own t
    inst16[1] = inst(rows(1, -4.5))
    out, rc = O.raytrace_tlas(tlas8, [0, 1], inst16, [[0, 0, 0, 0]], leaf_blas(), TRI,
                              [[0, 1, 2]], [0], ray((0, 0, 0), (0, 0, 1)))
    assert rc == 0 and out[0].tolist() == [np.float32(4.5).view(np.uint32), 0, 1, 0]


def test_blas_root_box_is_not_tested():
    # raytraceBlas pushes the BLAS root without a box test (scene.slang:172):
    # a root box the ray misses still has its triangle tested
    _, inst16 = _two_instances(8.0)
    far = np.stack([node([10, 10, 10], 0, [11, 11, 11], 0)])
    tlas8 = np.stack([node([-1, -1, 3], 0, [1, 1, 5], 0)])
    out, rc = O.raytrace_tlas(tlas8, [0], inst16[:1], [[0, 0, 0, 0]], far, TRI, [[0, 1, 2]], [0],
                              ray((0, 0, 0), (0, 0, 1)))
    assert rc == 0 and out[0].tolist() == [np.float32(5.0).view(np.uint32), 0, 0, 0]


def _chain(depth):
    """Node 2k+2 is internal with children (2k+3, 2k+4), every box the same,
    so each level leaves one sibling on the stack."""
    blas8 = [node([-1, -1, 0], 0, [1, 1, 0], 1)]
    for k in range(depth):
        blas8.append(node([-1, -1, 0], 0, [1, 1, 0], 0))
        blas8.append(node([-1, -1, 0], 0, [1, 1, 0], 0 if k == depth - 1 else 2 * k + 3))
    return np.stack(blas8)


@pytest.mark.parametrize("depth,full", [(30, False), (70, True)])
def test_stack_overflow_is_reported(depth, full):
    tlas8 = np.stack([node([-1, -1, 5], 0, [1, 1, 5], 0)])
    out, rc = O.raytrace_tlas(tlas8, [0], np.stack([inst(rows(1, -5))]), [[0, 0, 0, 0]],
                              _chain(depth), TRI, [[0, 1, 2]], [0], ray((0, 0, 0), (0, 0, 1)))
    assert rc == (1 if full else 0)
    if not full:
        assert out[0].tolist() == [np.float32(5.0).view(np.uint32), 0, 0, 0]


def _oracle_scene(seed, n_inst, scale_rot):
    rng = np.random.default_rng(seed)
    meshes = [scenes.soup(400, seed=seed + 1, extent=0.1),
              scenes.heightfield(12, 10, seed=seed + 2),
              scenes.clusters(300, k=4, seed=seed + 3), scenes.soup(3, seed=seed + 4, extent=0.3)]
    pos, tris_orig, tris_perm, blas_perm, blas8, blas4 = [], [], [], [], [], []
    vo = to = no = 0
    lo, hi = [], []
    for p, t in meshes:
        b = O.build(O.tri_aabbs(p, t))
        enc = O.encode(b["nodes"])
        lst = b["list"]
        pos.append(p)
        tris_orig.append(t)
        tris_perm.append(t[lst])
        blas_perm.append(lst)
        blas8.append(enc)
        blas4.append([no, to, vo, 0])
        bf = enc[0].view(np.float32)
        lo.append(bf[0:3])
        hi.append(bf[4:7])
        vo, to, no = vo + p.shape[0], to + t.shape[0], no + enc.shape[0]
    T = np.zeros((n_inst, 4, 4))
    if scale_rot:
        scale = rng.uniform(0.5, 2, n_inst)[:, None, None]
        T[:, :3, :3] = scenes._random_rotations(rng, n_inst) * scale
    else:
        T[:, :3, :3] = np.eye(3)
    T[:, :3, 3] = rng.uniform(-1.5, 1.5, (n_inst, 3)).astype(np.float32)
    T[:, 3, 3] = 1
    which = rng.integers(0, len(meshes), n_inst)
    inv = np.linalg.inv(T)[:, :3, :4].reshape(-1, 12).astype(np.float32)
    inst16 = np.stack([inst(inv[i].reshape(3, 4), which[i]) for i in range(n_inst)])
    boxes = O.instance_boxes(np.array(lo)[which], np.array(hi)[which],
                             T[:, :3, :].reshape(-1, 12).astype(np.float32))
    tl = O.build(boxes)
    mesh3 = np.array([[b[1], m[1].shape[0], b[2]] for b, m in zip(blas4, meshes)], np.uint32)
    return dict(tlas8=O.encode(tl["nodes"]), tlas_perm=tl["list"], inst16=inst16,
                blas4=np.array(blas4, np.uint32), blas8=np.concatenate(blas8),
                pos=np.concatenate(pos), tris_perm=np.concatenate(tris_perm),
                blas_perm=np.concatenate(blas_perm), tris_orig=np.concatenate(tris_orig),
                mesh3=mesh3, centers=T[:, :3, 3])


def _aimed_rays(rng, centers, n):
    tgt = centers[rng.integers(0, len(centers), n)] + rng.normal(0, 0.3, (n, 3))
    o = tgt + rng.normal(0, 1, (n, 3)) * 4
    d = (tgt - o).astype(np.float32)
    d /= np.sqrt((d.astype(np.float64) ** 2).sum(1))[:, None].astype(np.float32)
    return np.concatenate([o, d], 1).astype(np.float32)


def test_unscaled_scene_agrees_with_brute_force():
    s = _oracle_scene(5, 30, scale_rot=False)
    rng = np.random.default_rng(6)
    r = _aimed_rays(rng, s["centers"], 4000)
    d = r[:, 3:]
    unit = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]) == np.float32(1)
    r = r[unit]  # length(localRay.dir) == 1 exactly: no distance scaling
    assert r.shape[0] > 1000
    out, rc = O.raytrace_tlas(s["tlas8"], s["tlas_perm"], s["inst16"], s["blas4"], s["blas8"],
                              s["pos"], s["tris_perm"], s["blas_perm"], r)
    assert rc == 0
    t_ref, p_ref, i_ref = O.closest_hit_scene_brute(s["pos"], s["tris_orig"], s["mesh3"],
                                                    s["inst16"], r)
    t = out[:, 0].view(np.float32)
    hit = out[:, 1] != NONE
    assert (p_ref != NONE).sum() > r.shape[0] // 3
    # every reported hit is the reported triangle's own hit at that distance
    for k in np.flatnonzero(hit):
        rec = s["inst16"][out[k, 2]]
        m = rec[:12].view(np.float32).reshape(3, 4)
        b = s["blas4"][rec[12]]
        o = ((m[:, 0] * r[k, 0] + m[:, 1] * r[k, 1]) + m[:, 2] * r[k, 2]) + m[:, 3]
        dd = ((m[:, 0] * r[k, 3] + m[:, 1] * r[k, 4]) + m[:, 2] * r[k, 5]) + m[:, 3] * np.float32(0)
        tri = s["tris_orig"][b[1] + out[k, 1]] + b[2]
        assert out[k, 3] == b[1] + np.flatnonzero(s["blas_perm"][b[1]:b[1] + s["mesh3"][rec[12], 1]]
                                                  == out[k, 1])[0]
        tt = O.ray_tri(np.concatenate([o, dd]).astype(np.float32), *s["pos"][tri])
        assert np.float32(tt).view(np.uint32) == out[k, 0], k
    # never closer than the closest hit, and the same hit on nearly every ray
    both = hit & (p_ref != NONE)
    assert np.all(t[both] >= t_ref[both])
    assert not np.any(hit & (p_ref == NONE))
    same = (hit == (p_ref != NONE)) & (~hit | ((t == t_ref) & (out[:, 1] == p_ref)
                                               & (out[:, 2] == i_ref)))
    assert same.mean() > 0.995, same.mean()


def test_scaled_scene_hits_are_real_triangles():
    s = _oracle_scene(8, 40, scale_rot=True)
    rng = np.random.default_rng(9)
    r = _aimed_rays(rng, s["centers"], 3000)
    out, rc = O.raytrace_tlas(s["tlas8"], s["tlas_perm"], s["inst16"], s["blas4"], s["blas8"],
                              s["pos"], s["tris_perm"], s["blas_perm"], r)
    assert rc == 0
    hit = out[:, 1] != NONE
    t_ref, p_ref, _ = O.closest_hit_scene_brute(s["pos"], s["tris_orig"], s["mesh3"],
                                                s["inst16"], r)
    assert hit.sum() > 500
    assert not np.any(hit & (p_ref == NONE))  # a hit needs some triangle on the ray
    # the reported slot holds the reported triangle
    assert np.array_equal(s["blas_perm"][out[hit, 3]], out[hit, 1])
    assert np.all(out[~hit, 3] == NONE)
