"""glTF ingest (marigold_amd/gltf.py) against hand-built files with known
contents, following gltf_loading/src/lib.rs:26-220 and the glTF 2.0
specification (the reference's `gltf` crate is not vendored, so these known
answers are what pins the loader)."""
import base64
import json
import logging
import os

import numpy as np
import pytest

from marigold_amd import gltf


def _doc(buffers, views, accessors, meshes, nodes, scenes):
    return {"asset": {"version": "2.0"}, "buffers": buffers, "bufferViews": views,
            "accessors": accessors, "meshes": meshes, "nodes": nodes, "scenes": scenes}


def _write(tmp_path, doc, blobs):
    for name, data in blobs.items():
        (tmp_path / name).write_bytes(data)
    (tmp_path / "s.gltf").write_text(json.dumps(doc))
    return str(tmp_path)


def test_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    v0 = rng.standard_normal((7, 8)).astype(np.float32)
    t0 = np.array([[0, 1, 2], [2, 3, 4], [4, 5, 6]], np.uint32)
    v1 = rng.standard_normal((3, 8)).astype(np.float32)
    t1 = np.array([[0, 1, 2]], np.uint32)
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = [1.5, -2.25, 3.0]
    T2 = rng.standard_normal((4, 4)).astype(np.float32)
    T2[3] = [0, 0, 0, 1]
    gltf.write(str(tmp_path), [(v0, t0), (v1, t1)], [("a", [(T, 1), (T2, 0)]), ("b", [])])
    meshes, scenes = gltf.load(str(tmp_path))
    assert [m.key for m in meshes] == [(0, 0), (1, 0)]
    assert np.array_equal(meshes[0].vertices, v0) and np.array_equal(meshes[0].triangles, t0)
    assert np.array_equal(meshes[1].vertices, v1) and np.array_equal(meshes[1].triangles, t1)
    assert np.array_equal(meshes[0].bounds[0], v0[:, :3].min(0))
    assert np.array_equal(meshes[0].bounds[1], v0[:, :3].max(0))
    assert [s.name for s in scenes] == ["a", "b"]
    assert [i.mesh for i in scenes[0].instances] == [1, 0]
    assert np.array_equal(scenes[0].instances[0].transform, T)
    assert np.array_equal(scenes[0].instances[1].transform, T2)
    assert scenes[1].instances == []


def _known_file(tmp_path, drop=None):
    # interleaved POSITION + NORMAL (byteStride 24), u8-normalized UVs, u16
    # indices with a remainder of 2, one more normal than positions, a LINES
    # primitive, a sparse accessor, a data: buffer
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    nrm = np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0], [1, 1, 1], [9, 9, 9]], np.float32)
    inter = np.zeros((5, 6), np.float32)
    inter[:4, :3] = pos
    inter[:, 3:] = nrm
    uv = np.array([[0, 255], [51, 102], [255, 0], [1, 2]], np.uint8)
    idx = np.array([0, 1, 2, 1, 2, 3, 3, 0], np.uint16)
    sp_idx = np.array([2], np.uint8)
    sp_val = np.array([7], np.uint16)
    blob = (inter.tobytes() + uv.tobytes() + idx.tobytes() + sp_idx.tobytes() + b"\0"
            + sp_val.tobytes())
    o_uv, o_idx = 120, 128
    o_spi, o_spv = 144, 146
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": 120, "byteStride": 24},
             {"buffer": 0, "byteOffset": o_uv, "byteLength": 8},
             {"buffer": 0, "byteOffset": o_idx, "byteLength": 16},
             {"buffer": 0, "byteOffset": o_spi, "byteLength": 1},
             {"buffer": 0, "byteOffset": o_spv, "byteLength": 2},
             {"buffer": 1, "byteOffset": 0, "byteLength": 12}]
    acc = [{"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3",
            "min": [0, 0, 0], "max": [1, 1, 1.5]},
           {"bufferView": 0, "byteOffset": 12, "componentType": 5126, "count": 5, "type": "VEC3"},
           {"bufferView": 1, "componentType": 5121, "normalized": True, "count": 4, "type": "VEC2"},
           {"bufferView": 2, "componentType": 5123, "count": 8, "type": "SCALAR"},
           {"bufferView": 2, "componentType": 5123, "count": 6, "type": "SCALAR",
            "sparse": {"count": 1, "indices": {"bufferView": 3, "componentType": 5121},
                       "values": {"bufferView": 4}}},
           {"bufferView": 5, "componentType": 5126, "count": 1, "type": "VEC3",
            "min": [5, 5, 5], "max": [5, 5, 5]}]
    attrs = {"POSITION": 0, "NORMAL": 1, "TEXCOORD_0": 2}
    prims = [{"attributes": dict(attrs), "indices": 3},
             {"attributes": dict(attrs), "indices": 3, "mode": 1},
             {"attributes": dict(attrs), "indices": 4, "mode": 4}]
    if drop == "indices":
        del prims[0]["indices"]
    elif drop:
        del prims[0]["attributes"][drop]
    nodes = [{"translation": [1, 2, 3], "children": [1]},
             {"rotation": [0, 0, 1, 0], "scale": [2, 2, 2], "mesh": 0},
             {"mesh": 0, "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 4, 5, 6, 1]},
             {"mesh": 0, "rotation": [0, 0, 0.5 ** 0.5, 0.5 ** 0.5], "scale": [2, 3, 4]}]
    scenes = [{"nodes": [0]}, {"name": "second", "nodes": [2, 0]}, {"nodes": [3]}]
    doc = _doc([{"uri": "k.bin", "byteLength": len(blob)},
                {"uri": "data:application/octet-stream;base64,"
                 + base64.b64encode(np.zeros(3, np.float32).tobytes()).decode(),
                 "byteLength": 12}], views, acc,
               [{"primitives": prims}], nodes, scenes)
    return _write(tmp_path, doc, {"k.bin": blob})


def test_known_answers(tmp_path):
    meshes, scenes = gltf.load(_known_file(tmp_path))
    assert [m.key for m in meshes] == [(0, 0), (0, 2)]  # the LINES primitive is skipped
    m = meshes[0]
    assert m.vertices.shape == (4, 8)  # min(4 positions, 5 normals, 4 uvs)
    assert np.array_equal(m.vertices[:, 0:3], [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    assert np.array_equal(m.vertices[:, 4:7], [[0, 0, 1], [0, 1, 0], [1, 0, 0], [1, 1, 1]])
    u = np.array([[0, 255], [51, 102], [255, 0], [1, 2]], np.float32) / np.float32(255)
    assert np.array_equal(m.vertices[:, 3], u[:, 0]) and np.array_equal(m.vertices[:, 7], u[:, 1])
    assert m.triangles.tolist() == [[0, 1, 2], [1, 2, 3]]  # the last two indices dropped
    assert m.bounds[1].tolist() == [1, 1, 1.5]  # the accessor's max, not the vertices'
    assert meshes[1].triangles.tolist() == [[0, 1, 7], [1, 2, 3]]  # sparse substitution
    assert scenes[0].name == "unnamed scene #0" and scenes[1].name == "second"
    # T(1,2,3) * R(180 deg about z) * S(2)
    want = np.array([[-2, 0, 0, 1], [0, -2, 0, 2], [0, 0, 2, 3], [0, 0, 0, 1]], np.float32)
    assert [i.mesh for i in scenes[0].instances] == [0, 1]
    assert np.array_equal(scenes[0].instances[0].transform, want)
    moved = np.eye(4, dtype=np.float32)
    moved[:3, 3] = [4, 5, 6]
    assert np.array_equal(scenes[1].instances[0].transform, moved)  # column-major matrix
    assert np.array_equal(scenes[1].instances[2].transform, want)
    # R(90 deg about z) * S(2, 3, 4), not S * R
    rs = np.array([[0, -3, 0, 0], [2, 0, 0, 0], [0, 0, 4, 0], [0, 0, 0, 1]], np.float32)
    assert np.allclose(scenes[2].instances[0].transform, rs, atol=1e-6)


@pytest.mark.parametrize("drop,msg", [("POSITION", "positions"), ("NORMAL", "normals"),
                                      ("TEXCOORD_0", "TEXCOORD_0"), ("indices", "indices")])
def test_missing_attributes_fail(tmp_path, drop, msg):
    with pytest.raises(gltf.GltfError, match=msg):
        gltf.load(_known_file(tmp_path, drop=drop))


def test_no_gltf_and_short_view(tmp_path):
    with pytest.raises(gltf.GltfError, match="wasn't valid"):
        gltf.load(str(tmp_path))
    d = tmp_path / "x"
    d.mkdir()
    path = _known_file(d)
    doc = json.loads((d / "s.gltf").read_text())
    doc["accessors"][3]["count"] = 9  # one u16 past its 16-byte view
    (d / "s.gltf").write_text(json.dumps(doc))
    with pytest.raises(gltf.GltfError, match="past buffer view"):
        gltf.load(path)


def test_non_uniform_scale_warns(tmp_path, caplog):
    v = np.zeros((3, 8), np.float32)
    v[1, 0] = v[2, 1] = 1
    T = np.diag([1, 2, 1, 1]).astype(np.float32)
    gltf.write(str(tmp_path), [(v, np.array([[0, 1, 2]], np.uint32))], [("s", [(T, 0)])])
    with caplog.at_level(logging.WARNING, logger="marigold_amd.gltf"):
        gltf.load(str(tmp_path))
    assert any("non-uniform scale" in r.message for r in caplog.records)
    assert os.path.exists(tmp_path / "scene.bin")
