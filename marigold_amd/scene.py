"""Two-level scene on the device: one BLAS per mesh, a TLAS over instances.

Mirrors the scene upload of marigold (model.rs:137-230: meshes concatenated
with vertex/triangle offsets and a BLAS root per mesh; instance world boxes
from the 8 transformed corners of the mesh box; geometry/mod.rs:179-199: the
TLAS built over the instances with TLAS_MAX_DEPTH = 32, scene/mod.rs:28-29).
Every build and trace runs through the C ABI of libmgbvh.so.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from . import scenes

BLAS_MAX_DEPTH = 32
TLAS_MAX_DEPTH = 32


def world_to_object(transforms: np.ndarray) -> np.ndarray:
    """(M,4,4) object-to-world -> (M,12) float32 rows of the inverse's top 3x4."""
    inv = np.linalg.inv(transforms.astype(np.float64))
    return np.ascontiguousarray(inv[:, :3, :4].reshape(-1, 12).astype(np.float32))


class GpuScene:
    """Builds every mesh's BLAS and the TLAS on the device and keeps the
    concatenated arrays resident for tracing."""

    def __init__(self, ctx: N.Context, meshes, transforms: np.ndarray, instance_mesh: np.ndarray,
                 blas_max_depth: int = BLAS_MAX_DEPTH, tlas_max_depth: int = TLAS_MAX_DEPTH,
                 method: int = N.METHOD_PLAIN):
        self.ctx = ctx
        nm = len(meshes)
        self.n_meshes = nm
        self.n_instances = int(len(instance_mesh))
        vert_off, tri_off, node_off = [], [], []
        pos_all, tris_orig, tris_perm, perm_all, nodes_all, lens_all = [], [], [], [], [], []
        nv = nt = nn = 0
        mesh_min = np.empty((nm, 3), np.float32)
        mesh_max = np.empty((nm, 3), np.float32)
        self.blas_stats = []
        for p, t in meshes:
            t_work = np.ascontiguousarray(t, dtype=np.uint32).copy()
            nodes, lens, perm, _, st = ctx.build_tris(
                np.ascontiguousarray(p, dtype=np.float32), t_work,
                N.default_settings(max_depth=blas_max_depth, method=method))
            self.blas_stats.append(st)
            vert_off.append(nv); tri_off.append(nt); node_off.append(nn)
            mesh_min[len(vert_off) - 1] = nodes[0]["min"]
            mesh_max[len(vert_off) - 1] = nodes[0]["max"]
            pos_all.append(p); tris_orig.append(t); tris_perm.append(t_work)
            perm_all.append(perm); nodes_all.append(nodes); lens_all.append(lens)
            nv += p.shape[0]; nt += t.shape[0]; nn += nodes.shape[0]
        self.positions = np.ascontiguousarray(np.concatenate(pos_all).astype(np.float32))
        self.tris_orig = np.ascontiguousarray(np.concatenate(tris_orig).astype(np.uint32))
        self.tris_perm = np.ascontiguousarray(np.concatenate(tris_perm))
        self.blas_perm = np.ascontiguousarray(np.concatenate(perm_all))
        self.blas_nodes = np.ascontiguousarray(np.concatenate(nodes_all))
        self.blas_lens = np.ascontiguousarray(np.concatenate(lens_all))
        self.blas = np.zeros(nm, dtype=N.BLAS_DTYPE)
        self.blas["node_offset"] = node_off
        self.blas["tri_offset"] = tri_off
        self.blas["vertex_offset"] = vert_off
        self.mesh3 = np.stack([np.array(tri_off), [m[1].shape[0] for m in meshes],
                               np.array(vert_off)], axis=1).astype(np.uint32)
        self.transforms = np.ascontiguousarray(transforms, dtype=np.float32)
        self.instances = np.zeros(self.n_instances, dtype=N.INSTANCE_DTYPE)
        self.instances["world_to_object"] = world_to_object(self.transforms)
        self.instances["blas"] = np.asarray(instance_mesh, dtype=np.uint32)
        which = np.asarray(instance_mesh, dtype=np.int64)
        self.instance_boxes = scenes.instance_boxes(mesh_min[which], mesh_max[which],
                                                    self.transforms)
        tnodes, tlens, tperm, _, tst = ctx.build_aabbs(
            self.instance_boxes, N.default_settings(max_depth=tlas_max_depth, method=method,
                                                    name="tlas"))
        self.tlas_nodes, self.tlas_lens, self.tlas_perm, self.tlas_stats = tnodes, tlens, tperm, tst
        d = ctx.to_device
        self.d = {
            "pos": d(self.positions), "tris_perm": d(self.tris_perm),
            "blas_perm": d(self.blas_perm), "blas_nodes": d(self.blas_nodes),
            "blas_lens": d(self.blas_lens), "blas": d(self.blas),
            "instances": d(self.instances), "tlas_nodes": d(self.tlas_nodes),
            "tlas_lens": d(self.tlas_lens), "tlas_perm": d(self.tlas_perm),
        }

    @property
    def inst16(self) -> np.ndarray:
        """Instance records as (I,16) u32 words (the oracle's format)."""
        return self.instances.view(np.uint32).reshape(-1, 16)

    def trace(self, rays_dev: N.DeviceBuffer, nrays: int, hits_dev: N.DeviceBuffer):
        d = self.d
        self.ctx.trace_scene(d["tlas_nodes"], d["tlas_lens"], d["tlas_perm"], d["instances"],
                             self.n_instances, d["blas"], self.n_meshes, d["blas_nodes"],
                             d["blas_lens"], d["pos"], 12, d["tris_perm"], d["blas_perm"],
                             rays_dev, nrays, hits_dev)

    def bounce(self, rays_in, hits, nrays, seed, rays_out):
        d = self.d
        self.ctx.gen_bounce(rays_in, hits, nrays, d["instances"], self.n_instances, d["blas"],
                            self.n_meshes, d["pos"], 12, d["tris_perm"], seed, rays_out)


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
