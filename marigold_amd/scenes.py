"""Seeded synthetic meshes and scenes for the BVH builder workloads.

All generators are deterministic (numpy PCG64 with a fixed seed), return
float32 vertex positions (V, 3) and uint32 triangles (N, 3), and never emit
NaN or -0.0.

* ``soup``          N independent triangles: centroid ~ U[0,1]^3, each vertex
                    = centroid + U[-0.01, 0.01]^3 (65,536-triangle CPU workload).
* ``heightfield``   W x H cells over [0,1]^2, 2 triangles per cell with a fixed
                    diagonal, height = 0.2 * seeded 6-octave value noise
                    (persistence 0.5); y is up.  1024 x 512 gives the
                    1,048,576-triangle GPU workload.
* ``batched_scene`` 2,000 meshes with log-uniform triangle counts in
                    [100, 200,000], half soups and half heightfields, plus
                    10,000 instances (random rotation, scale U[0.5, 2],
                    translation U[-100, 100]^3).
* ``clusters``      N triangles in k Gaussian clusters (sigma U[0.01, 0.1]).
"""
from __future__ import annotations

import numpy as np

__all__ = ["soup", "heightfield", "value_noise", "batched_scene", "clusters", "random_transforms"]


def _canon(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float32)
    x[x == 0] = np.float32(0.0)  # -0.0 -> +0.0
    return x


def soup(n: int, seed: int = 1, extent: float = 0.01):
    """n independent triangles; vertex k of triangle t is 3t + k."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.random((n, 1, 3), dtype=np.float32)
    off = (rng.random((n, 3, 3), dtype=np.float32) * np.float32(2.0) - np.float32(1.0)) * np.float32(extent)
    pos = _canon((c + off).reshape(-1, 3))
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return np.ascontiguousarray(pos), tris


def value_noise(x: np.ndarray, y: np.ndarray, seed: int, octaves: int = 6, persistence: float = 0.5,
                base_freq: float = 4.0) -> np.ndarray:
    """Sum of `octaves` bilinear (smoothstep) value-noise layers in [-1, 1], float64."""
    total = np.zeros_like(x, dtype=np.float64)
    norm = 0.0
    amp = 1.0
    freq = base_freq
    for o in range(octaves):
        rng = np.random.Generator(np.random.PCG64(seed * 1009 + o))
        gx = int(np.ceil(freq)) + 2
        lattice = rng.random((gx + 1, gx + 1)) * 2.0 - 1.0
        u = x * freq
        v = y * freq
        i0 = np.floor(u).astype(np.int64)
        j0 = np.floor(v).astype(np.int64)
        fu = u - i0
        fv = v - j0
        su = fu * fu * (3.0 - 2.0 * fu)
        sv = fv * fv * (3.0 - 2.0 * fv)
        a = lattice[i0, j0]
        b = lattice[i0 + 1, j0]
        c = lattice[i0, j0 + 1]
        d = lattice[i0 + 1, j0 + 1]
        total += amp * ((a * (1 - su) + b * su) * (1 - sv) + (c * (1 - su) + d * su) * sv)
        norm += amp
        amp *= persistence
        freq *= 2.0
    return total / norm


def heightfield(w: int = 1024, h: int = 512, seed: int = 2, amplitude: float = 0.2):
    """(w+1) x (h+1) vertex grid over [0,1]^2, 2 triangles per cell."""
    xs = np.arange(w + 1, dtype=np.float64) / w
    zs = np.arange(h + 1, dtype=np.float64) / h
    X, Z = np.meshgrid(xs, zs, indexing="xy")  # (h+1, w+1)
    Y = amplitude * value_noise(X, Z, seed)
    pos = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    pos = _canon(pos.astype(np.float32))
    jj, ii = np.meshgrid(np.arange(h, dtype=np.uint32), np.arange(w, dtype=np.uint32), indexing="ij")
    v00 = (jj * (w + 1) + ii).ravel()
    v10 = v00 + 1
    v01 = v00 + (w + 1)
    v11 = v01 + 1
    t0 = np.stack([v00, v10, v11], axis=-1)
    t1 = np.stack([v00, v11, v01], axis=-1)
    tris = np.stack([t0, t1], axis=1).reshape(-1, 3).astype(np.uint32)
    return np.ascontiguousarray(pos), np.ascontiguousarray(tris)


def clusters(n: int, k: int = 64, seed: int = 4, tri_extent: float = 0.005):
    """n triangles around k Gaussian cluster centres (sigma ~ U[0.01, 0.1])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.random((k, 3))
    sig = rng.uniform(0.01, 0.1, size=k)
    which = rng.integers(0, k, size=n)
    c = centres[which] + rng.standard_normal((n, 3)) * sig[which, None]
    off = (rng.random((n, 3, 3)) * 2.0 - 1.0) * tri_extent
    pos = _canon((c[:, None, :] + off).reshape(-1, 3).astype(np.float32))
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return np.ascontiguousarray(pos), tris


def _random_rotations(rng, m):
    q = rng.standard_normal((m, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((m, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def random_transforms(rng, m: int, scale=(0.5, 2.0), translate=100.0) -> np.ndarray:
    """(m,4,4) float32 object-to-world: uniform random rotation (unit
    quaternion), uniform scale, uniform translation in [-translate, translate]^3."""
    R = _random_rotations(rng, m)
    s = rng.uniform(scale[0], scale[1], size=m)
    t = rng.uniform(-translate, translate, size=(m, 3))
    T = np.zeros((m, 4, 4))
    T[:, :3, :3] = R * s[:, None, None]
    T[:, :3, 3] = t
    T[:, 3, 3] = 1.0
    return T.astype(np.float32)


def batched_scene(n_meshes: int = 2000, n_instances: int = 10000, lo: int = 100, hi: int = 200_000,
                  seed: int = 3):
    """Meshes with log-uniform triangle counts (soup or heightfield) and
    instances of them.  Returns dict with lists `positions`, `tris`, `kinds`,
    `transforms` (M,4,4 float32 object-to-world) and `instance_mesh` (M,)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = np.exp(rng.uniform(np.log(lo), np.log(hi), size=n_meshes)).astype(np.int64)
    kinds = rng.integers(0, 2, size=n_meshes)  # 0 soup, 1 heightfield
    positions, tris = [], []
    for i in range(n_meshes):
        n = int(counts[i])
        if kinds[i] == 0:
            p, t = soup(n, seed=10_000 + i)
        else:
            cells = max(1, n // 2)
            w = max(1, int(round(np.sqrt(cells * 2))))
            h = max(1, cells // w)
            p, t = heightfield(w, h, seed=20_000 + i)
        positions.append(p)
        tris.append(t)
    which = rng.integers(0, n_meshes, size=n_instances)
    return {
        "positions": positions,
        "tris": tris,
        "kinds": kinds,
        "transforms": random_transforms(rng, n_instances),
        "instance_mesh": which.astype(np.uint32),
    }
