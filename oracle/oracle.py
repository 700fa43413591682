"""ctypes binding of the CPU oracle (oracle/bvh_oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or the CPU baseline.  It restates marigold's
`bvh/src/lib.rs` in C; see the header of bvh_oracle.c for the file:line map.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")
SRC = [os.path.join(HERE, "bvh_oracle.c"), os.path.join(HERE, "rays_oracle.c")]

NODE_DTYPE = np.dtype([("min", "<f4", (3,)), ("max", "<f4", (3,)), ("start", "<u4"), ("len", "<u4"),
                       ("child", "<u4")])
REC_DTYPE = np.dtype([("axis_method", "<u4"), ("value", "<u4"), ("count_lt", "<u4"),
                      ("cost_bits", "<u4")])


class Settings(C.Structure):
    _fields_ = [("min_leaf", C.c_uint32), ("max_leaf", C.c_uint32), ("max_depth", C.c_uint32),
                ("has_bounds", C.c_int32), ("bounds", C.c_float * 6)]


class Stats(C.Structure):
    _fields_ = [(k, C.c_uint32) for k in (
        "nnodes", "height", "n_binned", "n_adaptive", "n_median", "n_cutoff_leaves", "n_tie_events",
        "n_median_quirk", "n_assert_fail", "n_cost_leaf", "n_refused", "n_tie_benign")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build_lib(force: bool = False, plant: int = 0) -> str:
    """Compiles the oracle with gcc (-O3, no FMA contraction, no fast-math).
    plant != 0 builds one of the deliberately wrong variants (OR_PLANT in
    bvh_oracle.c) into its own library, for tests that show a check fails."""
    so = SO if not plant else os.path.join(HERE, f"liboracle_plant{plant}.so")
    if force or not os.path.exists(so) or any(os.path.getmtime(s) > os.path.getmtime(so) for s in SRC):
        cmd = ["gcc", "-O3", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
               "-pthread", f"-DOR_PLANT={int(plant)}", "-o", so] + SRC + ["-lm"]
        subprocess.run(cmd, check=True)
    return so


_lib = None
_planted = {}


def planted_lib(plant: int):
    """A deliberately wrong variant of the oracle (see OR_PLANT in bvh_oracle.c)."""
    if plant not in _planted:
        _planted[plant] = _bind(C.CDLL(build_lib(plant=plant)))
    return _planted[plant]


def lib():
    global _lib
    if _lib is None:
        build_lib()
        _lib = _bind(C.CDLL(SO))
    return _lib


def _bind(L):
    P = C.c_void_p
    if True:
        L.or_tri_aabbs.argtypes = [P, C.c_uint32, P, C.c_uint64, P]
        L.or_build.argtypes = [P, C.c_uint32, C.POINTER(Settings), P, P, P, C.POINTER(Stats)]
        L.or_build.restype = C.c_int
        L.or_encode_nodes.argtypes = [P, C.c_uint32, P]
        L.or_tree_sah.argtypes = [P, C.c_uint32]
        L.or_tree_sah.restype = C.c_double
        L.or_surface_area.argtypes = [P]
        L.or_surface_area.restype = C.c_float
        L.or_total_cmp_key.argtypes = [C.c_float]
        L.or_total_cmp_key.restype = C.c_int32
        L.or_bin_index.argtypes = [C.c_float, C.c_float, C.c_float, C.c_uint32]
        L.or_bin_index.restype = C.c_uint32
        L.or_binned_cost.argtypes = [P, P, P, C.c_uint32, C.c_uint32]
        L.or_binned_cost.restype = C.c_float
        L.or_closest_hit_brute.argtypes = [P, C.c_uint32, P, C.c_uint32, P, C.c_uint32, P, P]
        L.or_closest_hit_bvh.argtypes = [P, P, P, C.c_uint32, P, P, P, C.c_uint32, P, P]
        L.or_closest_hit_scene_brute.argtypes = [P, C.c_uint32, P, P, P, C.c_uint32, P,
                                                 C.c_uint32, P, P, P]
        L.or_raytrace_tlas.argtypes = [P, P, P, C.c_uint32, P, C.c_uint32, P, P, C.c_uint32, P,
                                       P, P, C.c_uint32, P]
        L.or_raytrace_tlas.restype = C.c_int
        L.or_wide8_collapse.argtypes = [P, P, C.c_uint32, C.c_float, C.c_float, C.c_uint32, P, P,
                                        P, P]
        L.or_wide8_collapse.restype = C.c_int
        L.or_trace_wide8.argtypes = [P, P, P, C.c_uint32, P, P, P, C.c_uint32, P, P]
        L.or_ray_tri.argtypes = [P, P, P, P]
        L.or_ray_tri.restype = C.c_float
        L.or_pcg.argtypes = [C.POINTER(C.c_uint32)]
        L.or_pcg.restype = C.c_uint32
        L.or_pcg_float.argtypes = [C.POINTER(C.c_uint32)]
        L.or_pcg_float.restype = C.c_float
        L.or_frame_jitter.argtypes = [C.c_uint32, C.c_int, P]
        L.or_gen_view_rays.argtypes = [P, P, C.c_uint32, C.c_uint32, C.c_float, C.c_float, P]
        L.or_gen_bounce.argtypes = [P, P, C.c_uint32, P, P, P, C.c_uint32, P, C.c_uint32, P]
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def tri_aabbs(pos: np.ndarray, tris: np.ndarray) -> np.ndarray:
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    out = np.empty((tris.shape[0], 6), dtype=np.float32)
    lib().or_tri_aabbs(_p(pos), pos.shape[1], _p(tris), tris.shape[0], _p(out))
    return out


def build(aabbs: np.ndarray, max_depth: int = 32, bounds=None, min_leaf: int = 1, max_leaf: int = 1,
          list_in: np.ndarray | None = None, L=None):
    """Runs BoundingVolumeHierarchy::new over objects with the given boxes.
    Returns dict(nodes, list, recs, stats)."""
    aabbs = np.ascontiguousarray(aabbs, dtype=np.float32)
    n = aabbs.shape[0]
    s = Settings(min_leaf, max_leaf, max_depth, 0)
    if bounds is not None:
        s.has_bounds = 1
        for a in range(3):
            s.bounds[a] = float(bounds[0][a])
            s.bounds[3 + a] = float(bounds[1][a])
    lst = np.arange(n, dtype=np.uint32) if list_in is None else np.array(list_in, dtype=np.uint32)
    cap = max(2 * n - 1, 1)
    nodes = np.zeros(cap, dtype=NODE_DTYPE)
    recs = np.zeros(cap, dtype=REC_DTYPE)
    st = Stats()
    rc = (L or lib()).or_build(_p(aabbs), n, C.byref(s), _p(lst), _p(nodes), _p(recs), C.byref(st))
    if rc != 0:
        raise RuntimeError(f"oracle build failed: {rc}")
    k = st.nnodes
    return {"nodes": nodes[:k], "list": lst, "recs": recs[:k], "stats": st.as_dict()}


def encode(nodes: np.ndarray) -> np.ndarray:
    """32-byte GPU records as (k, 8) uint32."""
    nodes = np.ascontiguousarray(nodes)
    out = np.empty((nodes.shape[0], 8), dtype=np.uint32)
    lib().or_encode_nodes(_p(nodes), nodes.shape[0], _p(out))
    return out


def tree_sah(nodes: np.ndarray, L=None) -> float:
    nodes = np.ascontiguousarray(nodes)
    return float((L or lib()).or_tree_sah(_p(nodes), nodes.shape[0]))


def surface_area(b6) -> float:
    a = np.ascontiguousarray(b6, dtype=np.float32)
    return float(lib().or_surface_area(_p(a)))


def total_cmp_key(f: float) -> int:
    return int(lib().or_total_cmp_key(C.c_float(f)))


def bin_index(c, cmin, cext, nbins=32) -> int:
    return int(lib().or_bin_index(C.c_float(c), C.c_float(cmin), C.c_float(cext), nbins))


def binned_cost(parent6, lt6, gt6, n_lt, n_gt) -> float:
    p = np.ascontiguousarray(parent6, dtype=np.float32)
    l = np.ascontiguousarray(lt6, dtype=np.float32)
    g = np.ascontiguousarray(gt6, dtype=np.float32)
    return float(lib().or_binned_cost(_p(p), _p(l), _p(g), n_lt, n_gt))


def ray_tri(ray6, v1, v2, v3) -> float:
    arrs = [np.ascontiguousarray(x, dtype=np.float32) for x in (ray6, v1, v2, v3)]
    return float(lib().or_ray_tri(*[_p(a) for a in arrs]))


def closest_hit_brute(pos, tris, rays):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    t = np.empty(rays.shape[0], dtype=np.float32)
    prim = np.empty(rays.shape[0], dtype=np.uint32)
    lib().or_closest_hit_brute(_p(pos), pos.shape[1], _p(tris), tris.shape[0], _p(rays),
                               rays.shape[0], _p(t), _p(prim))
    return t, prim


def closest_hit_bvh(nodes8, lens, pos, tris_perm, perm, rays, L=None):
    """Closest hit through a built tree: nodes8 = (k, 8) u32 node records.
    L: a planted variant's library (planted_lib) instead of the oracle's."""
    nodes8 = np.ascontiguousarray(nodes8, dtype=np.uint32)
    lens = None if lens is None else np.ascontiguousarray(lens, dtype=np.uint32)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    tris_perm = np.ascontiguousarray(tris_perm, dtype=np.uint32)
    perm = np.ascontiguousarray(perm, dtype=np.uint32)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    t = np.empty(rays.shape[0], dtype=np.float32)
    prim = np.empty(rays.shape[0], dtype=np.uint32)
    (L or lib()).or_closest_hit_bvh(_p(nodes8), _p(lens), _p(pos), pos.shape[1], _p(tris_perm),
                                    _p(perm), _p(rays), rays.shape[0], _p(t), _p(prim))
    return t, prim


def closest_hit_scene_brute(pos, tris, mesh3, inst16, rays):
    """Two-level closest hit by brute force: tris = concatenated original-order
    triangles, mesh3 = (M, 3) {first triangle, count, vertex offset},
    inst16 = (I, 16) u32 instance records (12 floats, mesh at word 12),
    rays = (N, 6).  Returns (t, prim, inst)."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    tris = np.ascontiguousarray(tris, dtype=np.uint32)
    mesh3 = np.ascontiguousarray(mesh3, dtype=np.uint32)
    inst16 = np.ascontiguousarray(inst16, dtype=np.uint32)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    n = rays.shape[0]
    t = np.empty(n, dtype=np.float32)
    prim = np.empty(n, dtype=np.uint32)
    inst = np.empty(n, dtype=np.uint32)
    lib().or_closest_hit_scene_brute(_p(pos), pos.shape[1], _p(tris), _p(mesh3), _p(inst16),
                                     inst16.shape[0], _p(rays), n, _p(t), _p(prim), _p(inst))
    return t, prim, inst


def raytrace_tlas(tlas8, tlas_perm, inst16, blas4, blas8, pos, tris_perm, blas_perm, rays):
    """The path tracer's own traversal (scene.slang:91-221), tree dependent.
    tlas8/blas8: (k, 8) u32 nodes; tlas_perm: TLAS start_index -> instance;
    inst16: (I, 16) u32 instance records; blas4: (M, 4) u32 {node offset,
    triangle offset, vertex offset, 0}; rays: (N, 6).  Returns ((N, 4) u32
    {distance bits, prim, inst, slot}, status bits)."""
    arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in
            (tlas8, tlas_perm, inst16, blas4, blas8, tris_perm, blas_perm)]
    tlas8, tlas_perm, inst16, blas4, blas8, tris_perm, blas_perm = arrs
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    out = np.empty((rays.shape[0], 4), dtype=np.uint32)
    rc = lib().or_raytrace_tlas(_p(tlas8), _p(tlas_perm), _p(inst16), inst16.shape[0], _p(blas4),
                                blas4.shape[0], _p(blas8), _p(pos), pos.shape[1], _p(tris_perm),
                                _p(blas_perm), _p(rays), rays.shape[0], _p(out))
    return out, int(rc)


WIDE8_DEFAULTS = {"c_node": 1.0, "c_prim": 0.3, "p_max": 3}


def wide8_collapse(nodes8, lens=None, c_node=1.0, c_prim=0.3, p_max=3):
    """Compressed 8-wide tree from binary nodes ((k, 8) u32).  Returns
    (wide (m, 80) u8, tri_index (objects,) u32, root cost)."""
    nodes8 = np.ascontiguousarray(nodes8, dtype=np.uint32)
    lens = None if lens is None else np.ascontiguousarray(lens, dtype=np.uint32)
    k = nodes8.shape[0]
    nobj = int(lens[nodes8[:, 7] == 0].sum()) if lens is not None else (k + 1) // 2
    out = np.zeros((max(k, 1), 80), dtype=np.uint8)
    tri = np.zeros(max(nobj, 1), dtype=np.uint32)
    nw = C.c_uint32()
    cost = C.c_float()
    rc = lib().or_wide8_collapse(_p(nodes8), _p(lens), k, c_node, c_prim, p_max, _p(out), _p(tri),
                                 C.byref(nw), C.byref(cost))
    if rc != 0:
        raise ValueError(f"wide8 collapse failed: {rc}")
    return out[:nw.value], tri[:nobj], float(cost.value)


def trace_wide8(wide, tri_index, pos, tris_perm, perm, rays, L=None):
    wide = np.ascontiguousarray(wide, dtype=np.uint8)
    tri_index = np.ascontiguousarray(tri_index, dtype=np.uint32)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    tris_perm = np.ascontiguousarray(tris_perm, dtype=np.uint32)
    perm = np.ascontiguousarray(perm, dtype=np.uint32)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    t = np.empty(rays.shape[0], dtype=np.float32)
    prim = np.empty(rays.shape[0], dtype=np.uint32)
    (L or lib()).or_trace_wide8(_p(wide), _p(tri_index), _p(pos), pos.shape[1], _p(tris_perm),
                                _p(perm), _p(rays), rays.shape[0], _p(t), _p(prim))
    return t, prim


def instance_boxes(mesh_min, mesh_max, o2w12):
    """World boxes of instances as model.rs:189-221 computes them: the 8
    corners of the mesh box through the transform (each row
    ((m0*x + m1*y) + m2*z) + m3 in float32, no FMA), then min and max.
    o2w12 = (M, 12) rows of the 3x4 object-to-world.  Returns (M, 6) float32."""
    lo = np.asarray(mesh_min, dtype=np.float32)
    hi = np.asarray(mesh_max, dtype=np.float32)
    m = np.asarray(o2w12, dtype=np.float32).reshape(-1, 3, 4)
    out = np.empty((m.shape[0], 6), dtype=np.float32)
    out[:, :3] = np.inf
    out[:, 3:] = -np.inf
    for k in range(8):
        x = hi[:, 0] if k & 4 else lo[:, 0]
        y = hi[:, 1] if k & 2 else lo[:, 1]
        z = hi[:, 2] if k & 1 else lo[:, 2]
        for a in range(3):
            w = m[:, a, 0] * x
            w = w + m[:, a, 1] * y
            w = w + m[:, a, 2] * z
            w = w + m[:, a, 3]
            out[:, a] = np.minimum(out[:, a], w)
            out[:, 3 + a] = np.maximum(out[:, 3 + a], w)
    return out


# ---------------------------------------------------------------- rays (rays_oracle.c)
def pcg_stream(seed: int, count: int, L=None):
    """count successive getUint() values of Random(seed) (random.slang:5-22)."""
    st = C.c_uint32(seed)
    return np.array([(L or lib()).or_pcg(C.byref(st)) for _ in range(count)], dtype=np.uint32)


def pcg_floats(seed: int, count: int, L=None):
    st = C.c_uint32(seed)
    return np.array([(L or lib()).or_pcg_float(C.byref(st)) for _ in range(count)],
                    dtype=np.float32)


def frame_jitter(frame_count: int, accumulate: bool, L=None):
    j = np.zeros(2, np.float32)
    (L or lib()).or_frame_jitter(frame_count, 1 if accumulate else 0, _p(j))
    return j


def gen_view_rays(inv_view_proj, position, width: int, height: int, frame_count: int = 0,
                  accumulate: bool = False, L=None):
    """(width * height, 8) float32 rays: org[3], 0, dir[3], +inf."""
    m = np.ascontiguousarray(np.asarray(inv_view_proj, np.float32).reshape(16))
    p = np.ascontiguousarray(np.asarray(position, np.float32).reshape(3))
    j = frame_jitter(frame_count, accumulate, L)
    out = np.empty((width * height, 8), np.float32)
    (L or lib()).or_gen_view_rays(_p(m), _p(p), width, height, float(j[0]), float(j[1]), _p(out))
    return out


def gen_bounce(rays8, hits4, pos, tris_perm, seed: int, inst16=None, blas_voff=None, L=None):
    """One bounce per ray: (n, 8) float32 rays from (n, 8) rays and (n, 4) u32 hits."""
    rays8 = np.ascontiguousarray(rays8, dtype=np.float32)
    hits4 = np.ascontiguousarray(hits4, dtype=np.uint32)
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    tris_perm = np.ascontiguousarray(tris_perm, dtype=np.uint32)
    inst16 = None if inst16 is None else np.ascontiguousarray(inst16, dtype=np.float32)
    blas_voff = None if blas_voff is None else np.ascontiguousarray(blas_voff, dtype=np.uint32)
    n = rays8.shape[0]
    out = np.empty((n, 8), np.float32)
    (L or lib()).or_gen_bounce(_p(rays8), _p(hits4), n, _p(inst16), _p(blas_voff), _p(pos),
                               pos.shape[1], _p(tris_perm), seed, _p(out))
    return out
