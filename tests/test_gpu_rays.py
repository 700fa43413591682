"""GPU ray generation against the oracle (oracle/rays_oracle.c), through the C
ABI: primary rays bit for bit (pathtrace.slang:46-66, screen.slang:48-72),
bounces to a few ulps (random.slang:29-41 uses sine and cosine, whose last
bits differ between the device's and the host's libm)."""
import numpy as np
import pytest

from marigold_amd import _native as N
from marigold_amd import scenes
from marigold_amd.scene import GpuScene, inverse_view_proj, look_at_camera
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rays8(r):
    return r.view(np.float32).reshape(-1, 8)


@pytest.mark.parametrize("cam,w,h,frame,acc", [
    (((0.5, 0.6, -0.2), (0.5, 0.0, 0.5), 0.6), 1024, 512, 0, False),
    (((0.5, 0.6, -0.2), (0.5, 0.0, 0.5), 0.6), 4096, 2048, 9, True),
    (((0.0, 60.0, -260.0), (0.0, 0.0, 0.0), 0.9), 640, 480, 31, True),
    (((3.0, -2.0, 5.0), (0.1, 0.2, 0.3), 1.3), 333, 77, 2, False),
])
def test_view_rays_bit_exact(ctx, cam, w, h, frame, acc):
    print(f"[case] view rays {w}x{h} frame {frame} accumulate {acc} camera {cam}", flush=True)
    eye, target, fov = cam
    m = inverse_view_proj(eye, target, fov, w / h)
    d = ctx.alloc(32 * w * h)
    ctx.gen_view(m, eye, w, h, d, frame_count=frame, accumulate=acc)
    got = _rays8(d.download(N.RAY_DTYPE, w * h))
    d.free()
    want = O.gen_view_rays(m, eye, w, h, frame_count=frame, accumulate=acc)
    bad = np.nonzero(np.any(got.view(np.uint32) != want.view(np.uint32), axis=1))[0]
    assert bad.size == 0, (bad.size, bad[:4], got[bad[:2]], want[bad[:2]])


def test_pinhole_rays_through_pixel_centres(ctx):
    print("[case] pinhole camera rays vs float64 pixel centres", flush=True)
    w, h = 400, 300
    cam = look_at_camera((0.5, 0.6, -0.2), (0.5, 0.0, 0.5), 0.6, w / h)
    d = ctx.alloc(32 * w * h)
    ctx.gen_primary(cam, w, h, d)
    r = d.download(N.RAY_DTYPE, w * h)
    d.free()
    fwd, right, up = (np.frombuffer(bytes(v), np.float32).astype(np.float64)
                      for v in (cam.fwd, cam.right, cam.up))
    th = np.tan(0.5 * cam.fov_y)
    px, py = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5, indexing="xy")
    want = (fwd + ((2 * px / w - 1) * th * cam.aspect).ravel()[:, None] * right
            + ((1 - 2 * py / h) * th).ravel()[:, None] * up)
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    assert np.all(r["org"] == np.float32([0.5, 0.6, -0.2]))
    assert np.abs(r["dir"] - want).max() < 2e-6
    assert np.all(r["reserved"] == 0) and np.all(np.isinf(r["tmax"]))


def _pcg_pair(seeds):
    """The first two getFloat() draws of Random(seed) for every seed (uint32 numpy;
    the same hash as random.slang:5-9, for the tolerance only)."""
    out = []
    s = seeds.astype(np.uint32).copy()
    for _ in range(2):
        with np.errstate(over="ignore"):
            state = s * np.uint32(747796405) + np.uint32(2891336453)
            word = ((state >> ((state >> np.uint32(28)) + np.uint32(4))) ^ state) * np.uint32(277803737)
        s = (word >> np.uint32(22)) ^ word
        out.append(s.astype(np.float32) / np.float32(4294967296.0))
    return out


def _cancel_scale(r8, hits4, pos, tris, seed, inst16=None, blas_voff=None):
    """|n + u| per hit ray in float64: the normalisation of n + u multiplies the
    few-ulp differences of sine and cosine by 1 / |n + u|."""
    n = r8.shape[0]
    u0, u1 = _pcg_pair(np.arange(n, dtype=np.uint64) + seed)
    phi, z = u0.astype(np.float64) * 2 * np.pi, 2 * u1.astype(np.float64) - 1
    rr = np.sqrt(np.maximum(0, 1 - z * z))
    u = np.stack([np.sin(phi) * rr, np.cos(phi) * rr, z], axis=1)
    hit = hits4[:, 1] != 0xFFFFFFFF
    slot = np.where(hit, hits4[:, 3], 0).astype(np.int64)  # a miss reads triangle 0
    inst = np.where(hit, hits4[:, 2], 0).astype(np.int64)
    voff = np.zeros(n, np.int64)
    if inst16 is not None:
        voff = blas_voff[inst16[inst, 12].view(np.uint32)].astype(np.int64)
        voff = np.where(hit, voff, 0)  # triangle 0 belongs to mesh 0 (vertex offset 0)
    tr = tris[slot].astype(np.int64) + voff[:, None]
    v = pos.astype(np.float64)[tr]
    nl = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    if inst16 is not None:
        m = inst16[inst, :12].astype(np.float64).reshape(-1, 3, 4)
        nl = np.einsum("kji,kj->ki", m[:, :, :3], nl)
    nl /= np.linalg.norm(nl, axis=1, keepdims=True)
    nl *= -np.sign(np.einsum("ki,ki->k", nl, r8[:, 4:7].astype(np.float64)))[:, None]
    return np.where(hit, np.linalg.norm(nl + u, axis=1), 1.0)


def _check_bounce(tag, got, want, scale):
    g, w = _rays8(got), want
    assert np.array_equal(g[:, [3, 7]].view(np.uint32), w[:, [3, 7]].view(np.uint32)), tag
    org = np.abs(g[:, :3] - w[:, :3]).max()
    derr = np.abs(g[:, 4:7] - w[:, 4:7]).max(axis=1)
    exact = np.mean(np.all(g.view(np.uint32) == w.view(np.uint32), axis=1))
    print(f"[bounce {tag}] max |org diff| {org:.3g}, max |dir diff| {derr.max():.3g}, "
          f"max |dir diff| x |n + u| {(derr * scale).max():.3g}, bit-identical rays {exact:.4f}",
          flush=True)
    # origins use no sine or cosine: the same operations, the same bits
    assert org == 0, tag
    bad = np.nonzero(derr > 1e-6 + 4e-7 / np.maximum(scale, 1e-3))[0]
    assert bad.size == 0, (tag, bad[:5], derr[bad[:5]], scale[bad[:5]])


def test_bounce_single_mesh(ctx):
    print("[case] heightfield 256x128, 512x256 primary rays, bounce vs oracle", flush=True)
    pos, tris = scenes.heightfield(256, 128, seed=5)
    t = tris.copy()
    nodes, lens, perm, _, _ = ctx.build_tris(pos, t, N.default_settings())
    d_nodes, d_lens, d_pos = ctx.to_device(nodes), ctx.to_device(lens), ctx.to_device(pos)
    d_tris, d_perm = ctx.to_device(t), ctx.to_device(perm)
    w, h = 512, 256
    n = w * h
    cam = look_at_camera((0.5, 0.6, -0.2), (0.5, 0.0, 0.5), 0.6, w / h)
    d_r0, d_r1, d_h0 = ctx.alloc(32 * n), ctx.alloc(32 * n), ctx.alloc(16 * n)
    ctx.gen_primary(cam, w, h, d_r0)
    ctx.trace_closest(d_nodes, d_lens, d_pos, 12, d_tris, d_perm, d_r0, n, d_h0)
    ctx.gen_bounce(d_r0, d_h0, n, None, 0, None, 0, d_pos, 12, d_tris, 4321, d_r1)
    r0 = d_r0.download(N.RAY_DTYPE, n)
    h0 = d_h0.download(N.HIT_DTYPE, n)
    got = d_r1.download(N.RAY_DTYPE, n)
    for b in (d_nodes, d_lens, d_pos, d_tris, d_perm, d_r0, d_r1, d_h0):
        b.free()
    hits = h0.view(np.uint32).reshape(-1, 4)
    assert n // 4 < (h0["prim"] != N.MISS).sum() < n  # hits and misses both occur
    want = O.gen_bounce(_rays8(r0), hits, pos, t, 4321)
    _check_bounce("single mesh", got, want, _cancel_scale(_rays8(r0), hits, pos, t, 4321))


def test_bounce_scene(ctx):
    print("[case] 4 meshes x 20 instances, 240x160 primary rays, bounce vs oracle", flush=True)
    rng = np.random.default_rng(8)
    meshes = [scenes.soup(500, seed=50, extent=0.1), scenes.heightfield(20, 16, seed=51),
              scenes.clusters(400, k=4, seed=52), scenes.soup(3, seed=53, extent=0.3)]
    k = 20
    R = scenes._random_rotations(rng, k)
    T = np.zeros((k, 4, 4))
    T[:, :3, :3] = R * rng.uniform(0.5, 2.0, k)[:, None, None]
    T[:, :3, 3] = rng.uniform(-1.5, 1.5, (k, 3))
    T[:, 3, 3] = 1.0
    sc = GpuScene(ctx, meshes, T.astype(np.float32), rng.integers(0, 4, k).astype(np.uint32))
    w, h = 240, 160
    n = w * h
    cam = look_at_camera((0.0, 1.0, -7.0), (0.0, 0.0, 0.0), 0.8, w / h)
    d_r0, d_r1, d_h0 = ctx.alloc(32 * n), ctx.alloc(32 * n), ctx.alloc(16 * n)
    ctx.gen_primary(cam, w, h, d_r0)
    sc.trace(d_r0, n, d_h0)
    sc.bounce(d_r0, d_h0, n, 77, d_r1)
    r0 = d_r0.download(N.RAY_DTYPE, n)
    h0 = d_h0.download(N.HIT_DTYPE, n)
    got = d_r1.download(N.RAY_DTYPE, n)
    tris_perm = sc.download("tris_perm", np.uint32, 3 * sc.n_tris).reshape(-1, 3)
    for b in (d_r0, d_r1, d_h0):
        b.free()
    assert (h0["prim"] != N.MISS).sum() > n // 20
    h4 = h0.view(np.uint32).reshape(-1, 4)
    i16 = sc.instances.view(np.float32).reshape(-1, 16)
    want = O.gen_bounce(_rays8(r0), h4, sc.positions, tris_perm, 77, inst16=i16,
                        blas_voff=sc.blas["vertex_offset"])
    scale = _cancel_scale(_rays8(r0), h4, sc.positions, tris_perm, 77, inst16=i16,
                          blas_voff=sc.blas["vertex_offset"])
    sc.free()
    _check_bounce("scene", got, want, scale)
