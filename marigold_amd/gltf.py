"""glTF 2.0 scene ingest (and a writer) for the builder's host side.

Follows marigold's loader, gltf_loading/src/lib.rs:26-220:

* the scene directory holds one ``.gltf`` file; its buffers are files next to
  it (``uri``), read whole (lib.rs:57-82);
* every primitive of mode TRIANGLES (4, the default) of every mesh becomes
  one mesh; others are skipped (lib.rs:86-91);
* vertices are MeshVertex records (mesh_interface/src/lib.rs:6-12: position,
  uv.x, normal, uv.y; 32 bytes) zipped from POSITION, NORMAL and TEXCOORD_0,
  so the count is the smallest of the three (lib.rs:98-124); texture
  coordinates stored as normalized u8/u16 are divided by 255/65535 in f32;
* triangles are the index accessor read as u32 in threes, a remainder
  dropped (lib.rs:127-133);
* the mesh bounds are the POSITION accessor's min/max (lib.rs:135-139), which
  the BLAS build takes as its root bounds (model.rs:83-84);
* instances come from each scene's node hierarchy, world = parent * local,
  with a warning on non-uniform scale (lib.rs:27-53); a scene without a
  name is "unnamed scene #i" (lib.rs:156-160).

The reference keeps meshes in a HashMap keyed by (mesh, primitive) and packs
them in its iteration order (lib.rs:190-200), which Rust randomises per
process; here the packed order is sorted by (mesh, primitive), so indices are
reproducible.  A missing POSITION, NORMAL, TEXCOORD_0 or index accessor is an
error, as the reference's `expect` calls are.  Sparse accessors and base64
``data:`` buffers are read as the glTF 2.0 specification defines them.
"""
from __future__ import annotations

import base64
import json
import logging
import os
from dataclasses import dataclass, field

import numpy as np

log = logging.getLogger(__name__)

_COMPONENT = {5120: "<i1", 5121: "<u1", 5122: "<i2", 5123: "<u2", 5125: "<u4", 5126: "<f4"}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}
_TRIANGLES = 4


class GltfError(ValueError):
    pass


@dataclass
class Mesh:
    """One triangle primitive: (V, 8) float32 MeshVertex records
    [px, py, pz, uv_x, nx, ny, nz, uv_y], (N, 3) uint32 triangles and the
    POSITION accessor's bounds."""
    vertices: np.ndarray
    triangles: np.ndarray
    bounds: tuple
    key: tuple  # (mesh index, primitive index) in the file

    @property
    def positions(self) -> np.ndarray:
        """(V, 3) float32 view with the 32-byte record stride."""
        return self.vertices[:, :3]


@dataclass
class Instance:
    transform: np.ndarray  # (4, 4) float32 object-to-world
    mesh: int              # packed mesh index


@dataclass
class Scene:
    name: str
    instances: list = field(default_factory=list)


def _find_gltf(path: str) -> str:
    names = sorted(n for n in os.listdir(path) if n.endswith(".gltf"))
    if not names:
        raise GltfError(f"gltf path {path} wasn't valid")
    return os.path.join(path, names[0])


def _load_buffers(doc: dict, base: str) -> list:
    out = []
    for b in doc.get("buffers", []):
        uri = b.get("uri")
        if uri is None:
            raise GltfError("a .gltf buffer without uri (binary chunks exist only in .glb)")
        if uri.startswith("data:"):
            data = base64.b64decode(uri.split(",", 1)[1])
        else:
            with open(os.path.join(base, uri), "rb") as f:
                data = f.read()
        if len(data) < b["byteLength"]:
            raise GltfError(f"buffer {uri[:40]} shorter than its byteLength")
        out.append(data)
    return out


def _view(doc, buffers, view_idx, byte_offset, dtype, ncomp, count):
    bv = doc["bufferViews"][view_idx]
    data = buffers[bv["buffer"]]
    elem = dtype.itemsize * ncomp
    stride = bv.get("byteStride") or elem
    start = bv.get("byteOffset", 0) + byte_offset
    end = start + (stride * (count - 1) + elem if count else 0)
    if end > bv.get("byteOffset", 0) + bv["byteLength"] or end > len(data):
        raise GltfError(f"accessor runs past buffer view {view_idx}")
    if count == 0:
        return np.zeros((0, ncomp), dtype=dtype)
    return np.ndarray((count, ncomp), dtype=dtype, buffer=data, offset=start,
                      strides=(stride, dtype.itemsize)).copy()


def read_accessor(doc: dict, buffers: list, idx: int) -> tuple:
    """(count, ncomp) array of accessor `idx` in its component type, with a
    sparse substitution applied; returns (array, accessor dict)."""
    acc = doc["accessors"][idx]
    dtype = np.dtype(_COMPONENT[acc["componentType"]])
    if acc["type"] not in _NCOMP:
        raise GltfError(f"accessor type {acc['type']} is not read here")
    ncomp = _NCOMP[acc["type"]]
    count = acc["count"]
    if "bufferView" in acc:
        arr = _view(doc, buffers, acc["bufferView"], acc.get("byteOffset", 0), dtype, ncomp, count)
    else:
        arr = np.zeros((count, ncomp), dtype=dtype)
    sp = acc.get("sparse")
    if sp:
        n = sp["count"]
        ii = sp["indices"]
        idt = np.dtype(_COMPONENT[ii["componentType"]])
        where = _view(doc, buffers, ii["bufferView"], ii.get("byteOffset", 0), idt, 1, n)[:, 0]
        vv = sp["values"]
        vals = _view(doc, buffers, vv["bufferView"], vv.get("byteOffset", 0), dtype, ncomp, n)
        if n and int(where.max()) >= count:
            raise GltfError("sparse index past the accessor's count")
        arr[where.astype(np.int64)] = vals
    return arr, acc


def _as_f32(arr: np.ndarray, acc: dict) -> np.ndarray:
    if arr.dtype == np.float32:
        return arr
    if not acc.get("normalized"):
        raise GltfError("texture coordinates must be float or normalized u8/u16")
    if arr.dtype == np.uint8:
        return arr.astype(np.float32) / np.float32(255.0)
    if arr.dtype == np.uint16:
        return arr.astype(np.float32) / np.float32(65535.0)
    raise GltfError(f"texture coordinates of type {arr.dtype} are not read")


def _read_primitive(doc, buffers, prim) -> tuple:
    attrs = prim.get("attributes", {})
    for name, msg in (("POSITION", "couldn't read mesh positions"),
                      ("NORMAL", "couldn't read mesh normals"),
                      ("TEXCOORD_0", "couldn't read mesh uvs from TEXCOORD_0")):
        if name not in attrs:
            raise GltfError(msg)
    if "indices" not in prim:
        raise GltfError("couldn't read mesh indices")
    pos, pacc = read_accessor(doc, buffers, attrs["POSITION"])
    nrm, _ = read_accessor(doc, buffers, attrs["NORMAL"])
    uv, uacc = read_accessor(doc, buffers, attrs["TEXCOORD_0"])
    if pos.dtype != np.float32 or nrm.dtype != np.float32 or pos.shape[1] != 3 or nrm.shape[1] != 3:
        raise GltfError("POSITION and NORMAL must be float VEC3")
    uv = _as_f32(uv, uacc)
    nv = min(pos.shape[0], nrm.shape[0], uv.shape[0])
    verts = np.empty((nv, 8), dtype=np.float32)
    verts[:, 0:3] = pos[:nv]
    verts[:, 3] = uv[:nv, 0]
    verts[:, 4:7] = nrm[:nv]
    verts[:, 7] = uv[:nv, 1]
    idx, iacc = read_accessor(doc, buffers, prim["indices"])
    if idx.dtype not in (np.uint8, np.uint16, np.uint32) or idx.shape[1] != 1:
        raise GltfError("indices must be unsigned SCALAR")
    idx = idx[:, 0].astype(np.uint32)
    tris = np.ascontiguousarray(idx[: idx.shape[0] // 3 * 3].reshape(-1, 3))
    if "min" not in pacc or "max" not in pacc:
        raise GltfError("POSITION accessor without min/max")
    bounds = (np.asarray(pacc["min"][:3], np.float32), np.asarray(pacc["max"][:3], np.float32))
    return verts, tris, bounds


def _quat_matrix(q) -> np.ndarray:
    x, y, z, w = (float(v) for v in q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def node_matrix(node: dict) -> np.ndarray:
    """Local transform of a node: `matrix` (column-major) or T * R * S,
    composed in double and rounded once to float32."""
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float32).reshape(4, 4).T.copy()
    t = node.get("translation", [0, 0, 0])
    r = node.get("rotation", [0, 0, 0, 1])
    s = node.get("scale", [1, 1, 1])
    m = np.eye(4)
    m[:3, :3] = _quat_matrix(r) * np.asarray(s, np.float64)[None, :]
    m[:3, 3] = t
    return m.astype(np.float32)


def mat_mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """float32 4x4 product, each entry summed left to right."""
    out = np.empty((4, 4), np.float32)
    for i in range(4):
        for j in range(4):
            acc = np.float32(a[i, 0] * b[0, j])
            for k in range(1, 4):
                acc = np.float32(acc + np.float32(a[i, k] * b[k, j]))
            out[i, j] = acc
    return out


def _uniform_scale(m: np.ndarray) -> bool:
    s = np.linalg.norm(m[:3, :3].astype(np.float64), axis=0)
    return bool(np.all(s == s[0]))


def _traverse(doc, ni, parent, out, depth=0):
    if depth > 1024:
        raise GltfError("node hierarchy deeper than 1024 (a cycle?)")
    node = doc["nodes"][ni]
    local = node_matrix(node)
    if not _uniform_scale(local):
        log.warning("Encountered a non-uniform scale in a gltf transform")
    world = mat_mul(parent, local)
    if "mesh" in node:
        for pi, prim in enumerate(doc["meshes"][node["mesh"]]["primitives"]):
            if prim.get("mode", _TRIANGLES) == _TRIANGLES:
                out.append((world, (node["mesh"], pi)))
    for c in node.get("children", []):
        _traverse(doc, c, world, out, depth + 1)


def load(path: str) -> tuple:
    """Loads the .gltf in directory `path`.  Returns (meshes, scenes): a list
    of Mesh in (mesh, primitive) order and a list of Scene whose instances
    index into it."""
    gpath = _find_gltf(path)
    with open(gpath, "r", encoding="utf-8") as f:
        doc = json.load(f)
    buffers = _load_buffers(doc, path)
    meshes = []
    for mi, m in enumerate(doc.get("meshes", [])):
        for pi, prim in enumerate(m.get("primitives", [])):
            if prim.get("mode", _TRIANGLES) != _TRIANGLES:
                continue
            verts, tris, bounds = _read_primitive(doc, buffers, prim)
            meshes.append(Mesh(verts, tris, bounds, (mi, pi)))
            log.info("Mesh #%s of mesh at path %s has %d vertices and %d triangles", (mi, pi),
                     path, verts.shape[0], tris.shape[0])
    packed = {m.key: i for i, m in enumerate(meshes)}
    scenes = []
    for si, sc in enumerate(doc.get("scenes", [])):
        found = []
        for ni in sc.get("nodes", []):
            _traverse(doc, ni, np.eye(4, dtype=np.float32), found)
        scenes.append(Scene(sc.get("name", f"unnamed scene #{si}"),
                            [Instance(w, packed[k]) for w, k in found]))
    log.info("Creating %d scenes (%d unique meshes)", len(scenes), len(meshes))
    return meshes, scenes


def write(path: str, meshes: list, scenes: list, name: str = "scene") -> str:
    """Writes meshes ((V, 8) MeshVertex records or (V, 3) positions, plus (N, 3)
    triangles) and scenes (list of (name, [(4x4 transform, mesh index)])) as
    `path/name.gltf` + `path/name.bin`, one glTF mesh per input mesh, u32
    indices, one node per instance with a matrix.  Returns the .gltf path."""
    os.makedirs(path, exist_ok=True)
    blob = bytearray()
    views, accessors, gmeshes = [], [], []

    def add(arr, target, typ, ctype, minmax=False):
        arr = np.ascontiguousarray(arr)
        while len(blob) % 4:
            blob.append(0)
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": arr.nbytes,
                      "target": target})
        blob.extend(arr.tobytes())
        acc = {"bufferView": len(views) - 1, "componentType": ctype, "count": int(arr.shape[0]),
               "type": typ}
        if minmax:
            acc["min"] = [float(v) for v in arr.min(axis=0)]
            acc["max"] = [float(v) for v in arr.max(axis=0)]
        accessors.append(acc)
        return len(accessors) - 1

    for verts, tris in meshes:
        verts = np.asarray(verts, np.float32)
        if verts.shape[1] == 3:
            full = np.zeros((verts.shape[0], 8), np.float32)
            full[:, 0:3] = verts
            full[:, 5] = 1.0  # normal +y
            verts = full
        p = add(verts[:, 0:3], 34962, "VEC3", 5126, minmax=True)
        n = add(verts[:, 4:7], 34962, "VEC3", 5126)
        u = add(verts[:, [3, 7]], 34962, "VEC2", 5126)
        i = add(np.asarray(tris, np.uint32).reshape(-1), 34963, "SCALAR", 5125)
        gmeshes.append({"primitives": [{"attributes": {"POSITION": p, "NORMAL": n,
                                                       "TEXCOORD_0": u}, "indices": i}]})
    nodes, gscenes = [], []
    for sname, insts in scenes:
        roots = []
        for T, mi in insts:
            nodes.append({"mesh": int(mi),
                          "matrix": [float(v) for v in np.asarray(T, np.float32).T.reshape(16)]})
            roots.append(len(nodes) - 1)
        gscenes.append({"name": sname, "nodes": roots})
    doc = {"asset": {"version": "2.0"}, "buffers": [{"uri": f"{name}.bin",
                                                     "byteLength": len(blob)}],
           "bufferViews": views, "accessors": accessors, "meshes": gmeshes, "nodes": nodes,
           "scenes": gscenes}
    with open(os.path.join(path, f"{name}.bin"), "wb") as f:
        f.write(bytes(blob))
    gpath = os.path.join(path, f"{name}.gltf")
    with open(gpath, "w", encoding="utf-8") as f:
        json.dump(doc, f)
    return gpath
