"""Two-level scene on the device: one BLAS per mesh, a TLAS over instances.

Mirrors the scene upload of marigold (model.rs:137-230: meshes concatenated
with vertex/triangle offsets and a BLAS root per mesh, instance world boxes
from the 8 transformed corners of the mesh box; geometry/mod.rs:179-199: the
TLAS built over the instances with TLAS_MAX_DEPTH = 32, scene/mod.rs:28-29).
Every build and trace runs on the device through the C ABI of libmgbvh.so:

* ``mode="plain"``   one mgbvh_build_tris_device call per mesh, as the
                     reference calls BoundingVolumeHierarchy::new per mesh;
* ``mode="batched"`` one mgbvh_build_blas_batch_device call for all meshes.

Both give identical node arrays (the batched build is bit-identical per mesh).
"""
from __future__ import annotations

import numpy as np

from . import _native as N

BLAS_MAX_DEPTH = 32
TLAS_MAX_DEPTH = 32


def world_to_object(transforms: np.ndarray) -> np.ndarray:
    """(M,4,4) object-to-world -> (M,12) float32 rows of the inverse's top 3x4."""
    inv = np.linalg.inv(transforms.astype(np.float64))
    return np.ascontiguousarray(inv[:, :3, :4].reshape(-1, 12).astype(np.float32))


class GpuScene:
    """Scene arrays resident on the device; `build()` rebuilds every BLAS,
    the instance boxes and the TLAS from them without host round trips of
    the geometry."""

    def __init__(self, ctx: N.Context, meshes, transforms: np.ndarray, instance_mesh: np.ndarray,
                 mode: str = "batched", blas_max_depth: int = BLAS_MAX_DEPTH,
                 tlas_max_depth: int = TLAS_MAX_DEPTH, build: bool = True,
                 method: int = N.METHOD_PLAIN, learned: dict | None = None):
        assert mode in ("plain", "batched")
        self.ctx = ctx
        self.mode = mode
        self.n_meshes = len(meshes)
        self.n_instances = int(len(instance_mesh))
        counts = np.array([t.shape[0] for _, t in meshes], dtype=np.int64)
        vcounts = np.array([p.shape[0] for p, _ in meshes], dtype=np.int64)
        self.tri_offsets = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
        self.vert_offsets = np.concatenate([[0], np.cumsum(vcounts)[:-1]]).astype(np.int64)
        slots = np.where(counts > 0, 2 * counts - 1, 0)
        self.node_offsets = np.concatenate([[0], np.cumsum(slots)[:-1]]).astype(np.int64)
        self.n_tris = int(counts.sum())
        self.n_slots = int(slots.sum())
        self.counts = counts
        self.positions = np.ascontiguousarray(np.concatenate([p for p, _ in meshes]).astype(np.float32))
        self.tris_orig = np.ascontiguousarray(np.concatenate([t for _, t in meshes]).astype(np.uint32))
        self.mesh_desc = np.zeros(self.n_meshes, dtype=N.MESH_DTYPE)
        self.mesh_desc["tri_offset"] = self.tri_offsets
        self.mesh_desc["tri_count"] = counts
        self.mesh_desc["vertex_offset"] = self.vert_offsets
        self.mesh_desc["vertex_count"] = vcounts
        self.mesh3 = np.stack([self.tri_offsets, counts, self.vert_offsets], axis=1).astype(np.uint32)
        self.blas = np.zeros(self.n_meshes, dtype=N.BLAS_DTYPE)
        self.blas["node_offset"] = self.node_offsets
        self.blas["tri_offset"] = self.tri_offsets
        self.blas["vertex_offset"] = self.vert_offsets
        self.transforms = np.ascontiguousarray(transforms, dtype=np.float32)
        self.o2w = np.ascontiguousarray(self.transforms[:, :3, :].reshape(-1, 12))
        self.instance_mesh = np.ascontiguousarray(instance_mesh, dtype=np.uint32)
        self.instances = np.zeros(self.n_instances, dtype=N.INSTANCE_DTYPE)
        self.instances["world_to_object"] = world_to_object(self.transforms)
        self.instances["blas"] = self.instance_mesh
        self.blas_settings = N.default_settings(max_depth=blas_max_depth, method=method,
                                                **(learned or {}))
        self.tlas_settings = N.default_settings(max_depth=tlas_max_depth, name="tlas")
        a, up = ctx.alloc, ctx.to_device
        ni = self.n_instances
        self.d = {
            "pos": up(self.positions), "tris": up(self.tris_orig),
            "tris_perm": a(12 * self.n_tris), "blas_perm": a(4 * self.n_tris),
            "blas_nodes": a(32 * self.n_slots), "blas_lens": a(4 * self.n_slots),
            "blas": up(self.blas), "instances": up(self.instances), "o2w": up(self.o2w),
            "inst_blas": up(self.instance_mesh), "boxes": a(24 * ni),
            "tlas_nodes": a(32 * max(2 * ni - 1, 1)), "tlas_lens": a(4 * max(2 * ni - 1, 1)),
            "tlas_perm": a(4 * max(ni, 1)),
        }
        self.n_tlas_nodes = 0
        if build:
            self.build()

    # ------------------------------------------------------------------ build
    def build_blas(self):
        c, d = self.ctx, self.d
        if self.mode == "batched":
            blas, counts, st = c.build_blas_batch(
                d["pos"], 12, self.positions.shape[0], d["tris"], self.mesh_desc,
                self.blas_settings, d["blas_nodes"], d["blas_lens"], d["blas_perm"],
                d["tris_perm"])
            assert np.array_equal(blas["node_offset"], self.blas["node_offset"])
            self.blas_node_counts = counts
        else:
            counts = np.zeros(self.n_meshes, dtype=np.uint32)
            for m in range(self.n_meshes):
                n = int(self.counts[m])
                if n == 0:
                    continue
                counts[m] = c.build_tris_device(
                    d["pos"], 12, int(self.mesh_desc[m]["vertex_count"]), d["tris"], n,
                    self.blas_settings, d["blas_nodes"], d["blas_lens"], d["blas_perm"],
                    d["tris_perm"], pos_offset=int(self.vert_offsets[m]),
                    tri_offset=int(self.tri_offsets[m]), node_offset=int(self.node_offsets[m]))
            self.blas_node_counts = counts

    def build_tlas(self):
        c, d = self.ctx, self.d
        c.instance_boxes(d["blas_nodes"], d["blas"], self.n_meshes, d["o2w"], d["inst_blas"],
                         self.n_instances, d["boxes"])
        self.n_tlas_nodes = c.build_aabbs_device(d["boxes"], self.n_instances, self.tlas_settings,
                                                 d["tlas_nodes"], d["tlas_lens"], d["tlas_perm"])

    def build(self):
        if getattr(self, "plan", None) is not None:
            ms, self.blas_node_counts, self.n_tlas_nodes = self.plan.launch()
            self.last_graph_ms = ms
            return
        self.build_blas()
        self.build_tlas()

    def record_graph(self):
        """Records the whole scene build (batched BLAS levels, instance boxes,
        TLAS) as one HIP graph; later `build()` calls replay it."""
        assert self.mode == "batched"
        if getattr(self, "plan", None) is not None:
            self.plan.close()
        d = self.d
        self.plan = N.ScenePlan(self.ctx, d["pos"], 12, self.positions.shape[0], d["tris"],
                                self.mesh_desc, self.blas_settings, d["blas_nodes"],
                                d["blas_lens"], d["blas_perm"], d["tris_perm"], d["o2w"],
                                d["inst_blas"], self.n_instances, self.tlas_settings, d["boxes"],
                                d["tlas_nodes"], d["tlas_lens"], d["tlas_perm"])

    def blas_sah(self) -> np.ndarray:
        """SAH cost of every BLAS (reference constants, relative to its root)."""
        return self.ctx.tree_sah(self.d["blas_nodes"], self.d["blas_lens"], self.blas,
                                 self.blas_node_counts)

    def scene_sah(self) -> float:
        """Triangle-weighted mean of the BLAS SAH costs."""
        return float((self.blas_sah() * self.counts).sum() / max(self.counts.sum(), 1))

    # ------------------------------------------------------------------ readback
    def download(self, name, dtype, count):
        return self.d[name].download(dtype, count)

    @property
    def blas_perm(self) -> np.ndarray:
        return self.download("blas_perm", np.uint32, self.n_tris)

    @property
    def inst16(self) -> np.ndarray:
        """Instance records as (I,16) u32 words (the oracle's format)."""
        return self.instances.view(np.uint32).reshape(-1, 16)

    # ------------------------------------------------------------------ trace
    def trace(self, rays_dev: N.DeviceBuffer, nrays: int, hits_dev: N.DeviceBuffer):
        d = self.d
        self.ctx.trace_scene(d["tlas_nodes"], d["tlas_lens"], d["tlas_perm"], d["instances"],
                             self.n_instances, d["blas"], self.n_meshes, d["blas_nodes"],
                             d["blas_lens"], d["pos"], 12, d["tris_perm"], d["blas_perm"],
                             rays_dev, nrays, hits_dev)

    def raytrace(self, rays_dev: N.DeviceBuffer, nrays: int, hits_dev: N.DeviceBuffer):
        """The path tracer's own traversal over these trees (tree dependent)."""
        d = self.d
        self.ctx.raytrace_tlas(d["tlas_nodes"], d["tlas_perm"], d["instances"], self.n_instances,
                               d["blas"], self.n_meshes, d["blas_nodes"], d["pos"], 12,
                               d["tris_perm"], d["blas_perm"], rays_dev, nrays, hits_dev)

    def bounce(self, rays_in, hits, nrays, seed, rays_out):
        d = self.d
        self.ctx.gen_bounce(rays_in, hits, nrays, d["instances"], self.n_instances, d["blas"],
                            self.n_meshes, d["pos"], 12, d["tris_perm"], seed, rays_out)

    def free(self):
        if getattr(self, "plan", None) is not None:
            self.plan.close()
            self.plan = None
        for b in self.d.values():
            b.free()
        self.d = {}


def inverse_view_proj(pos, target, fov_y: float, aspect: float, near: float = 0.1,
                      far: float = 100.0, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """inverse(perspective_rh(fov_y, aspect, near, far) * look_at_rh(pos,
    target, up)) as glam builds them (depth 0..1; camera.rs:105-110, 355-362),
    computed in float64 and returned as 16 float32 in glam's column-major order."""
    eye = np.asarray(pos, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    view[:3, 3] = [-s @ eye, -u @ eye, f @ eye]
    h = 1.0 / np.tan(0.5 * fov_y)
    r = far / (near - far)
    proj = np.zeros((4, 4))
    proj[0, 0], proj[1, 1] = h / aspect, h
    proj[2, 2], proj[2, 3], proj[3, 2] = r, r * near, -1.0
    inv = np.linalg.inv(proj @ view)
    return np.ascontiguousarray(inv.T.reshape(16).astype(np.float32))  # column-major


def look_at_camera(pos, target, fov_y: float, aspect: float, up=(0.0, 1.0, 0.0)) -> N.Camera:
    pos = np.asarray(pos, np.float64)
    f = np.asarray(target, np.float64) - pos
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    cam = N.Camera()
    for a in range(3):
        cam.pos[a] = pos[a]
        cam.fwd[a] = f[a]
        cam.right[a] = r[a]
        cam.up[a] = u[a]
    cam.fov_y = fov_y
    cam.aspect = aspect
    return cam
