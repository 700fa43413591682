"""Ray generation of the oracle (oracle/rays_oracle.c) pinned without itself:
published and hand-derived known answers, a float64 camera, distribution
invariants, and planted mistakes that the checks catch.  Reference:
marigold/assets/shaders/slang/lib/random.slang:5-41,
lib/screen.slang:48-72, pathtrace.slang:46-66."""
import numpy as np
import pytest

from marigold_amd.scene import inverse_view_proj
from oracle import oracle as O


def test_pcg_known_answer_and_uniformity():
    # pcg_hash(0) = 129708002 (Jarzynski and Olano, "Hash Functions for GPU
    # Rendering", JCGT 9(3), 2020, the hash random.slang:5-9 implements)
    assert int(O.pcg_stream(0, 1)[0]) == 129708002
    # the stream feeds each value back as the next seed
    s = O.pcg_stream(77, 5)
    for k in range(4):
        assert O.pcg_stream(int(s[k]), 1)[0] == s[k + 1]
    # getFloat: uniform on [0, 1] (64-bin chi-square over 200,000 draws)
    f = O.pcg_floats(12345, 200_000)
    assert f.min() >= 0.0 and f.max() <= 1.0
    hist = np.bincount(np.minimum((f * 64).astype(np.int64), 63), minlength=64)
    e = f.size / 64
    chi2 = ((hist - e) ** 2 / e).sum()
    assert chi2 < 120  # 63 degrees of freedom: p < 1e-5 above 120


def test_view_rays_known_answers():
    # M = I, camera at the origin: world = clip = (2 tx - 1, 1 - 2 ty, 1)
    # at texcoord = pixel / size; the top row (y = 0) looks up (+y)
    r = O.gen_view_rays(np.eye(4, dtype=np.float32).T.reshape(16), (0, 0, 0), 4, 2)
    assert np.all(r[:, :3] == 0) and np.all(r[:, 3] == 0) and np.all(np.isinf(r[:, 7]))
    d = r[:, 4:7].reshape(2, 4, 3).astype(np.float64)
    want = {(0, 0): (-1, 1, 1), (2, 0): (0, 1, 1), (1, 1): (-0.5, 0, 1), (2, 1): (0, 0, 1),
            (3, 1): (0.5, 0, 1)}
    for (x, y), v in want.items():
        v = np.array(v, np.float64) / np.linalg.norm(v)
        assert np.abs(d[y, x] - v).max() < 2.4e-7, (x, y, d[y, x], v)
    # a translation in M moves the far point, and the origin is the position
    m = np.eye(4)
    m[:3, 3] = (0.0, 0.0, 1.0)  # world = clip + (0, 0, 1)
    r = O.gen_view_rays(m.T.reshape(16).astype(np.float32), (0, 0, 0), 2, 2)
    assert np.abs(r[3, 4:7] - np.array([0, 0, 1], np.float32)).max() == 0  # pixel (1, 1)
    # w != 1: world = t.xyz / t.w (a row 3 of (0, 0, 0, 2) halves every point)
    m = np.eye(4)
    m[3, 3] = 2.0
    r2 = O.gen_view_rays(m.T.reshape(16).astype(np.float32), (0.25, 0, 0), 4, 2)
    v = np.array([-0.5, 0.5, 0.5]) - [0.25, 0, 0]  # pixel (0, 0): clip / 2 - position
    assert np.abs(r2[0, 4:7] - v / np.linalg.norm(v)).max() < 2.4e-7


def _camera_dirs(eye, target, fov, aspect, w, h, jitter):
    # float64: the direction through texcoord = (pixel + jitter) / size of a
    # look-at camera, from its basis (no matrices involved)
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, [0.0, 1.0, 0.0])
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    th = np.tan(0.5 * fov)
    px, py = np.meshgrid(np.arange(w) + float(jitter[0]), np.arange(h) + float(jitter[1]),
                         indexing="xy")
    cx = (2 * px / w - 1).ravel() * th * aspect
    cy = (1 - 2 * py / h).ravel() * th
    d = f + cx[:, None] * s + cy[:, None] * u
    return d / np.linalg.norm(d, axis=1, keepdims=True)


@pytest.mark.parametrize("frame,acc", [(0, False), (5, True), (123, True)])
def test_view_rays_of_a_camera(frame, acc):
    eye, target, fov, w, h = (0.5, 0.6, -0.2), (0.5, 0.0, 0.5), 0.6, 96, 48
    m = inverse_view_proj(eye, target, fov, w / h)
    r = O.gen_view_rays(m, eye, w, h, frame_count=frame, accumulate=acc)
    j = O.frame_jitter(frame, acc)
    assert np.all(np.abs(j) <= 0.5) and (np.any(j != 0) if acc else np.all(j == 0))
    want = _camera_dirs(eye, target, fov, w / h, w, h, j)
    assert np.all(r[:, :3] == np.float32(eye))
    assert np.abs(r[:, 4:7] - want).max() < 2e-6
    # planted: pixel centres instead of the jitter, and an unflipped y
    if acc:
        r7 = O.gen_view_rays(m, eye, w, h, frame, acc, L=O.planted_lib(7))
        assert np.abs(r7[:, 4:7] - want).max() > 1e-4
    r6 = O.gen_view_rays(m, eye, w, h, frame, acc, L=O.planted_lib(6))
    assert np.abs(r6[:, 4:7] - want).max() > 1e-2


def test_bounce_distribution_and_stream():
    # a unit quad in y = 0 (two triangles), rays straight down onto it
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1]], np.float32)
    tris = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    n = 200_000
    rng = np.random.default_rng(3)
    xz = rng.uniform(0.01, 0.99, (n, 2)).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2] = xz[:, 0], 2.0, xz[:, 1]
    rays[:, 5] = -1.0
    rays[:, 7] = np.inf
    hits = np.zeros((n, 4), np.uint32)
    hits[:, 0] = np.float32(2.0).view(np.uint32)
    hits[:, 1] = 0
    hits[:, 2] = 0
    hits[:, 3] = (np.arange(n) % 2).astype(np.uint32)
    hits[n // 2:, 1] = 0xFFFFFFFF  # the second half missed
    out = O.gen_bounce(rays, hits, pos, tris, seed=99)
    hit, miss = out[: n // 2], out[n // 2:]
    # origin: the hit point lifted 0.001 along the normal that faces the ray (+y)
    assert np.abs(hit[:, 1] - 0.001).max() < 1e-7
    assert np.array_equal(hit[:, [0, 2]], rays[: n // 2][:, [0, 2]])
    # cosine-weighted about +y: E[cos] = 2/3, E[cos^2] = 1/2 (6 sigma bounds)
    c = hit[:, 5].astype(np.float64)
    assert np.all(c > 0) and np.abs(np.linalg.norm(hit[:, 4:7], axis=1) - 1).max() < 1e-6
    assert abs(c.mean() - 2 / 3) < 6 * np.sqrt(1 / 18 / c.size)
    assert abs((c * c).mean() - 0.5) < 6 * np.sqrt((1 / 3 - 1 / 4) / c.size)
    # a miss keeps its origin and takes the unit vector itself: uniform on the
    # sphere, and exactly the stream's draws u0, u1 of Random(i + seed):
    # (sin 2 pi u0, cos 2 pi u0) sqrt(1 - z^2), z = 2 u1 - 1
    assert np.array_equal(miss[:, :3], rays[n // 2:, :3])
    d = miss[:, 4:7].astype(np.float64)
    assert np.abs(d.mean(axis=0)).max() < 6 * np.sqrt(1 / 3 / d.shape[0])
    assert abs((d[:, 2] ** 2).mean() - 1 / 3) < 6 * np.sqrt(4 / 45 / d.shape[0])
    for i in (n // 2, n // 2 + 1, n - 1):
        u0, u1 = O.pcg_floats(i + 99, 2).astype(np.float64)
        z = 2 * u1 - 1
        v = np.array([np.sin(2 * np.pi * u0), np.cos(2 * np.pi * u0)]) * np.sqrt(1 - z * z)
        assert np.abs(out[i, 4:7] - [v[0], v[1], z]).max() < 2e-6
