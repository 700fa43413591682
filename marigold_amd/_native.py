"""ctypes binding of the mgbvh C ABI (include/mgbvh.h).

Loads ``libmgbvh.so`` built in-tree next to this file and fails loudly when it
is missing: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmgbvh.so")

OK = 0
ERR_INVALID_ARG = -1
ERR_UNSUPPORTED = -2
ERR_NONFINITE = -3
ERR_INDEX = -4
ERR_DEVICE = -5
ERR_NO_DEVICE = -6
ERR_INTERNAL = -7

METHOD_PLAIN = 0
METHOD_LEARNED = 1

NODE_DTYPE = np.dtype([("min", "<f4", (3,)), ("start_index", "<u4"), ("max", "<f4", (3,)),
                       ("child_node", "<u4")])
SPLIT_DTYPE = np.dtype([("axis_method", "<u4"), ("value", "<u4"), ("count_lt", "<u4"),
                        ("cost_bits", "<u4")])
assert NODE_DTYPE.itemsize == 32 and SPLIT_DTYPE.itemsize == 16
RAY_DTYPE = np.dtype([("org", "<f4", (3,)), ("reserved", "<u4"), ("dir", "<f4", (3,)),
                      ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<u4"), ("inst", "<u4"), ("tri_slot", "<u4")])
BLAS_DTYPE = np.dtype([("node_offset", "<u4"), ("tri_offset", "<u4"), ("vertex_offset", "<u4"),
                       ("reserved", "<u4")])
INSTANCE_DTYPE = np.dtype([("world_to_object", "<f4", (12,)), ("blas", "<u4"),
                           ("reserved", "<u4", (3,))])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16
assert BLAS_DTYPE.itemsize == 16 and INSTANCE_DTYPE.itemsize == 64
MISS = 0xFFFFFFFF
MESH_DTYPE = np.dtype([("tri_offset", "<u4"), ("tri_count", "<u4"), ("vertex_offset", "<u4"),
                       ("vertex_count", "<u4")])
AABB_DTYPE = np.dtype([("min", "<f4", (3,)), ("max", "<f4", (3,))])


class Tuning(C.Structure):
    _fields_ = [("bin_blocks_per_cu", C.c_uint32), ("partition_blocks_per_cu", C.c_uint32),
                ("sweep_blocks_per_cu", C.c_uint32), ("small_blocks_per_cu", C.c_uint32),
                ("compute_units", C.c_uint32), ("partition_waves", C.c_uint32),
                ("bin_waves", C.c_uint32), ("sweep_pack_min_nodes", C.c_uint32),
                ("medium_max_objects", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("fov_y", C.c_float), ("aspect", C.c_float)]


def make_rays(org, dirs, tmax=np.inf) -> np.ndarray:
    """(N,3) origins and directions -> RAY_DTYPE array."""
    dirs = np.asarray(dirs, dtype=np.float32)
    r = np.zeros(dirs.shape[0], dtype=RAY_DTYPE)
    r["org"] = org
    r["dir"] = dirs
    r["tmax"] = tmax
    return r


class Settings(C.Structure):
    _fields_ = [
        ("name", C.c_char_p),
        ("has_bounds", C.c_int32),
        ("bounds_min", C.c_float * 3),
        ("bounds_max", C.c_float * 3),
        ("max_depth", C.c_uint32),
        ("profiling_info", C.c_int32),
        ("profiling_info_directory", C.c_char_p),
        ("min_leaf_objects", C.c_uint32),
        ("max_leaf_objects", C.c_uint32),
        ("method", C.c_uint32),
        ("learned_min_objects", C.c_uint32),
        ("learned_window", C.c_uint32),
        ("learned_guard", C.c_float),
        ("learned_no_network", C.c_uint32),
        ("reserved", C.c_uint32 * 3),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("construction_time", C.c_double),
        ("node_count", C.c_uint32),
        ("leaf_node_count", C.c_uint32),
        ("min_leaf_object_count", C.c_uint32),
        ("max_leaf_object_count", C.c_uint32),
        ("avg_leaf_object_count", C.c_double),
        ("stddev_leaf_object_count", C.c_double),
        ("total_object_count", C.c_uint32),
        ("height", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("n_large_levels", C.c_uint32),
        ("n_median", C.c_uint32),
        ("n_cutoff_leaves", C.c_uint32),
        ("n_guard_fallbacks", C.c_uint32),
        ("large_object_visits", C.c_uint64),
        ("large_nodes", C.c_uint32),
        ("n_window_splits", C.c_uint32),
        ("reserved", C.c_uint32 * 1),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class Wide8Params(C.Structure):
    _fields_ = [("c_node", C.c_float), ("c_prim", C.c_float), ("p_max", C.c_uint32),
                ("reserved", C.c_uint32)]


WIDE_NODE_BYTES = 80


class MgbvhError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mgbvh error {code}: {msg}")
        self.code = code


_lib = None

_P = C.c_void_p
_U32 = C.c_uint32
_U64 = C.c_uint64
_I = C.c_int


def _sig(lib):
    S = C.POINTER(Settings)
    St = C.POINTER(Stats)
    U64P = C.POINTER(C.c_uint64)
    lib.mgbvh_settings_default.argtypes = [S]
    lib.mgbvh_settings_default.restype = None
    lib.mgbvh_last_error.argtypes = []
    lib.mgbvh_last_error.restype = C.c_char_p
    lib.mgbvh_ctx_create.argtypes = [_I, C.POINTER(_P)]
    lib.mgbvh_ctx_destroy.argtypes = [_P]
    lib.mgbvh_ctx_stream.argtypes = [_P]
    lib.mgbvh_ctx_stream.restype = _P
    lib.mgbvh_ctx_synchronize.argtypes = [_P]
    lib.mgbvh_build_tris.argtypes = [_P, _P, _U32, _U64, _P, _U64, S, _P, _P, U64P, _P, _P, St]
    lib.mgbvh_build_aabbs.argtypes = [_P, _P, _U64, S, _P, _P, U64P, _P, _P, St]
    lib.mgbvh_build_tris_device.argtypes = [_P, _P, _U32, _U64, _P, _U64, S, _P, _P, _P, _P, _P,
                                            U64P, St]
    lib.mgbvh_build_aabbs_device.argtypes = [_P, _P, _U64, S, _P, _P, _P, _P, U64P, St]
    lib.mgbvh_device_alloc.argtypes = [_P, _U64, C.POINTER(_P)]
    lib.mgbvh_device_free.argtypes = [_P, _P]
    lib.mgbvh_copy_h2d.argtypes = [_P, _P, _P, _U64]
    lib.mgbvh_copy_d2h.argtypes = [_P, _P, _P, _U64]
    lib.mgbvh_event_record.argtypes = [_P, _I]
    lib.mgbvh_event_elapsed_ms.argtypes = [_P, _I, _I, C.POINTER(C.c_float)]
    lib.mgbvh_ctx_set_profiling.argtypes = [_P, _U32]
    lib.mgbvh_ctx_kernel_times.argtypes = [_P, _P, _P, _I]
    for fn in ("mgbvh_ctx_set_profiling", "mgbvh_ctx_kernel_times", "mgbvh_ctx_create", "mgbvh_ctx_destroy", "mgbvh_ctx_synchronize", "mgbvh_build_tris",
               "mgbvh_build_aabbs", "mgbvh_build_tris_device", "mgbvh_build_aabbs_device",
               "mgbvh_device_alloc", "mgbvh_device_free", "mgbvh_copy_h2d", "mgbvh_copy_d2h",
               "mgbvh_event_record", "mgbvh_event_elapsed_ms"):
        getattr(lib, fn).restype = _I
    _sig_extra(lib)


def _sig_extra(lib):
    """Signatures of entry points added beside the builder (traversal, batching)."""
    for name, (argtypes, restype) in _EXTRA.items():
        if hasattr(lib, name):
            f = getattr(lib, name)
            f.argtypes = argtypes
            f.restype = restype


_EXTRA: dict = {
    "mgbvh_trace_closest_device": ([_P, _P, _P, _P, _U32, _P, _P, _P, _U64, _P], _I),
    "mgbvh_trace_scene_device": ([_P, _P, _P, _P, _P, _U32, _P, _U32, _P, _P, _P, _U32, _P, _P,
                                  _P, _U64, _P], _I),
    "mgbvh_raytrace_tlas_device": ([_P, _P, _P, _P, _U32, _P, _U32, _P, _P, _U32, _P, _P, _P,
                                    _U64, _P], _I),
    "mgbvh_wide8_collapse_device": ([_P, _P, _P, _U64, _P, _P, _U64, _P, _U64, C.POINTER(_U64),
                                     C.POINTER(C.c_float)], _I),
    "mgbvh_wide8_collapse_batch_device": ([_P, _P, _P, _U64, _P, _P, _U32, _P, _P, _U64, _P, _U64,
                                           _P, _P, C.POINTER(_U64)], _I),
    "mgbvh_trace_wide8_device": ([_P, _P, _U64, _P, _P, _U32, _P, _P, _P, _U64, _P], _I),
    "mgbvh_gen_primary_rays_device": ([_P, C.POINTER(Camera), _U32, _U32, _P], _I),
    "mgbvh_gen_view_rays_device": ([_P, _P, _P, _U32, _U32, _U32, C.c_int, _P], _I),
    "mgbvh_copy_d2d": ([_P, _P, _P, _U64], _I),
    "mgbvh_build_tris_device_noalloc": ([_P, _P, _U32, _U64, _P, _U64, C.POINTER(Settings), _P, _P,
                                         _P, _P, _P, C.POINTER(_U64)], _I),
    "mgbvh_instance_boxes_device": ([_P, _P, _P, _U32, _P, _P, _U32, _P], _I),
    "mgbvh_model_create": ([_P, _U64, C.POINTER(_P)], _I),
    "mgbvh_model_destroy": ([_P], _I),
    "mgbvh_model_get_params": ([_P, _P, _P], _I),
    "mgbvh_model_set_params": ([_P, _P, _P], _I),
    "mgbvh_collect_samples_device": ([_P, _P, _U32, _U64, _P, _U64, C.POINTER(Settings), _U32,
                                      _P, _P, _U32, C.POINTER(_U32)], _I),
    "mgbvh_model_gradients_device": ([_P, _P, _P, _P, _U32, _P, C.POINTER(C.c_float)], _I),
    "mgbvh_model_train_device": ([_P, _P, _P, _P, _U32, _U32, _U32, C.c_float, _U64,
                                  C.POINTER(C.c_float), C.POINTER(C.c_float)], _I),
    "mgbvh_model_infer_device": ([_P, _P, _P, _U32, _P], _I),
    "mgbvh_ctx_set_model": ([_P, _P], _I),
    "mgbvh_scene_plan_create": ([_P, _P, _U32, _U64, _P, _P, _U32, C.POINTER(Settings), _P, _P,
                                 _P, _P, _P, _P, _U32, C.POINTER(Settings), _P, _P, _P, _P,
                                 C.POINTER(_P)], _I),
    "mgbvh_scene_plan_launch": ([_P, C.POINTER(C.c_float), _P, C.POINTER(C.c_uint64)], _I),
    "mgbvh_scene_plan_destroy": ([_P], _I),
    "mgbvh_scene_plan_info": ([_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                               C.POINTER(C.c_uint32)], _I),
    "mgbvh_ctx_set_tuning": ([_P, _P], _I),
    "mgbvh_ctx_get_tuning": ([_P, _P], _I),
    "mgbvh_tree_sah_device": ([_P, _P, _P, _P, _P, _U32, _P], _I),
    "mgbvh_build_blas_batch_device": ([_P, _P, _U32, _U64, _P, _P, _U32, _P, C.POINTER(Settings),
                                       _P, _P, _P, _P, _P, _P, _P, C.POINTER(Stats)], _I),
    "mgbvh_gen_bounce_rays_device": ([_P, _P, _P, _U64, _P, _U32, _P, _U32, _P, _U32, _P, _U32,
                                      _P], _I),
}


def lib():
    """The loaded libmgbvh.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                " (there is no CPU fallback)")
        _lib = C.CDLL(LIB_PATH)
        _sig(_lib)
    return _lib


def last_error() -> str:
    return lib().mgbvh_last_error().decode(errors="replace")


def check(rc: int):
    if rc != OK:
        raise MgbvhError(rc, last_error())


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def default_settings(max_depth: int = 32, bounds=None, method: int = METHOD_PLAIN, name: str = "",
                     profiling_info: bool = False, profiling_info_directory: str | None = None,
                     learned_min_objects: int = 0, learned_window: int = 0,
                     learned_guard: float = 0.0, learned_no_network: bool = False):
    s = Settings()
    lib().mgbvh_settings_default(C.byref(s))
    s.max_depth = max_depth
    s.method = method
    s.learned_min_objects = learned_min_objects
    s.learned_window = learned_window
    s.learned_guard = learned_guard
    s.learned_no_network = 1 if learned_no_network else 0
    s._name = name.encode()  # keep the bytes alive
    s.name = s._name
    s.profiling_info = 1 if profiling_info else 0
    if profiling_info_directory is not None:
        s._dir = str(profiling_info_directory).encode()
        s.profiling_info_directory = s._dir
    if bounds is not None:
        lo, hi = bounds
        s.has_bounds = 1
        for a in range(3):
            s.bounds_min[a] = float(lo[a])
            s.bounds_max[a] = float(hi[a])
    return s


class DeviceBuffer:
    """Device allocation owned by a Context (freed with it or by `free`)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().mgbvh_device_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p

    @property
    def value(self):
        return self.ptr.value

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().mgbvh_copy_h2d(self.ctx.handle, self.ptr, ptr(a), a.nbytes))

    def download(self, dtype, count, offset: int = 0) -> np.ndarray:
        """`count` elements of `dtype` starting at element `offset`."""
        out = np.empty(count, dtype=dtype)
        start = int(offset) * out.itemsize
        assert start + out.nbytes <= self.nbytes
        check(lib().mgbvh_copy_d2h(self.ctx.handle, ptr(out), C.c_void_p(self.ptr.value + start),
                                   out.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().mgbvh_device_free(self.ctx.handle, self.ptr)
        self.ptr = None


MODEL_FEATURES = 102
MODEL_CLASSES = 96
MODEL_PARAMS = 45408


class Model:
    """The learned build method's split proposer (device parameters)."""

    def __init__(self, ctx: "Context", seed: int = 1):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib().mgbvh_model_create(ctx.handle, seed, C.byref(h)))
        self.handle = h

    def params(self) -> np.ndarray:
        p = np.zeros(MODEL_PARAMS, dtype=np.float32)
        check(lib().mgbvh_model_get_params(self.ctx.handle, self.handle, ptr(p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (MODEL_PARAMS,)
        check(lib().mgbvh_model_set_params(self.ctx.handle, self.handle, ptr(p)))

    def gradients(self, feats, labels, n: int, grads=None):
        loss = C.c_float()
        check(lib().mgbvh_model_gradients_device(self.ctx.handle, self.handle, _dv(feats),
                                                 _dv(labels), n, _dv(grads), C.byref(loss)))
        return loss.value

    def train(self, feats, labels, n: int, epochs: int, batch: int = 256, lr: float = 1e-3,
              seed: int = 1):
        loss, acc = C.c_float(), C.c_float()
        check(lib().mgbvh_model_train_device(self.ctx.handle, self.handle, _dv(feats), _dv(labels),
                                             n, epochs, batch, lr, seed, C.byref(loss),
                                             C.byref(acc)))
        return loss.value, acc.value

    def infer(self, feats, n: int, logits):
        check(lib().mgbvh_model_infer_device(self.ctx.handle, self.handle, _dv(feats), n,
                                             _dv(logits)))

    def close(self):
        if self.handle is not None:
            lib().mgbvh_model_destroy(self.handle)
            self.handle = None


class ScenePlan:
    """A whole scene build recorded as one HIP graph (mgbvh_scene_plan_*)."""

    def __init__(self, ctx: "Context", pos, vstride, nverts, tris, meshes, blas_settings, blas_nodes,
                 blas_lens, blas_perm, tris_out, o2w, inst_blas, ninst, tlas_settings, boxes,
                 tlas_nodes, tlas_lens, tlas_perm):
        self.ctx = ctx
        self.meshes = np.ascontiguousarray(meshes, dtype=MESH_DTYPE)
        self._keep = (blas_settings, tlas_settings)
        h = C.c_void_p()
        check(lib().mgbvh_scene_plan_create(
            ctx.handle, _dv(pos), vstride, nverts, _dv(tris), ptr(self.meshes), len(self.meshes),
            C.byref(blas_settings), _dv(blas_nodes), _dv(blas_lens), _dv(blas_perm), _dv(tris_out),
            _dv(o2w), _dv(inst_blas), ninst, C.byref(tlas_settings) if tlas_settings else None,
            _dv(boxes), _dv(tlas_nodes), _dv(tlas_lens), _dv(tlas_perm), C.byref(h)))
        self.handle = h

    def launch(self):
        """Replays the graph; returns (device ms, BLAS node counts, TLAS node count)."""
        ms, nt = C.c_float(), C.c_uint64()
        counts = np.zeros(max(len(self.meshes), 1), dtype=np.uint32)
        check(lib().mgbvh_scene_plan_launch(self.handle, C.byref(ms), ptr(counts), C.byref(nt)))
        return ms.value, counts[:len(self.meshes)], nt.value

    def info(self) -> dict:
        """Levels the graph runs (BLAS, TLAS) and launches rebuilt level by level."""
        b, t, r = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().mgbvh_scene_plan_info(self.handle, C.byref(b), C.byref(t), C.byref(r)))
        return {"blas_levels": b.value, "tlas_levels": t.value, "launches_rebuilt": r.value}

    def close(self):
        if self.handle is not None:
            lib().mgbvh_scene_plan_destroy(self.handle)
            self.handle = None


def _dv(b):
    """Device pointer of a DeviceBuffer (None -> NULL)."""
    return None if b is None else b.ptr


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().mgbvh_ctx_create(device, C.byref(h)))
        self.handle = h
        self._bufs = []

    def alloc(self, nbytes: int) -> DeviceBuffer:
        b = DeviceBuffer(self, nbytes)
        self._bufs.append(b)
        return b

    def to_device(self, a: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(a)
        b = self.alloc(max(a.nbytes, 16))
        b.upload(a)
        return b

    def synchronize(self):
        check(lib().mgbvh_ctx_synchronize(self.handle))

    def record(self, slot: int):
        check(lib().mgbvh_event_record(self.handle, slot))

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        check(lib().mgbvh_event_elapsed_ms(self.handle, a, b, C.byref(ms)))
        return float(ms.value)

    KERNEL_KINDS = ["setup", "root", "bin", "sweep", "partition", "small", "compact", "permute",
                    "trace", "learn", "medium"]

    def set_tuning(self, **kw):
        t = self.get_tuning_struct()
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(lib().mgbvh_ctx_set_tuning(self.handle, C.byref(t)))

    def get_tuning_struct(self) -> "Tuning":
        t = Tuning()
        check(lib().mgbvh_ctx_get_tuning(self.handle, C.byref(t)))
        return t

    def get_tuning(self) -> dict:
        return self.get_tuning_struct().as_dict()

    model = None  # the network attached for the learned method

    def set_model(self, model):
        check(lib().mgbvh_ctx_set_model(self.handle, None if model is None else model.handle))
        self.model = model

    def set_profiling(self, kinds):
        mask = 0
        for k in kinds:
            mask |= 1 << self.KERNEL_KINDS.index(k)
        check(lib().mgbvh_ctx_set_profiling(self.handle, mask))

    def kernel_times(self, reset: bool = True):
        ms = np.zeros(16, dtype=np.float64)
        n = np.zeros(16, dtype=np.uint64)
        check(lib().mgbvh_ctx_kernel_times(self.handle, ptr(ms), ptr(n), 1 if reset else 0))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.KERNEL_KINDS) if n[i]}

    def close(self):
        if self.handle is not None:
            for b in self._bufs:
                b.free()
            self._bufs = []
            lib().mgbvh_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- traversal (device buffers: DeviceBuffer or None) ----------------------
    def trace_closest(self, nodes, lens, pos, vstride, tris_perm, perm, rays, nrays, hits):
        check(lib().mgbvh_trace_closest_device(self.handle, _dv(nodes), _dv(lens), _dv(pos),
                                               vstride, _dv(tris_perm), _dv(perm), _dv(rays),
                                               nrays, _dv(hits)))

    def trace_scene(self, tlas_nodes, tlas_lens, tlas_perm, instances, ninstances, blas, nblas,
                    blas_nodes, blas_lens, pos, vstride, tris_perm, blas_perm, rays, nrays, hits):
        check(lib().mgbvh_trace_scene_device(
            self.handle, _dv(tlas_nodes), _dv(tlas_lens), _dv(tlas_perm), _dv(instances),
            ninstances, _dv(blas), nblas, _dv(blas_nodes), _dv(blas_lens), _dv(pos), vstride,
            _dv(tris_perm), _dv(blas_perm), _dv(rays), nrays, _dv(hits)))

    def raytrace_tlas(self, tlas_nodes, tlas_perm, instances, ninstances, blas, nblas,
                      blas_nodes, pos, vstride, tris_perm, blas_perm, rays, nrays, hits):
        """The path tracer's own traversal (tree dependent, scene.slang:91-221)."""
        check(lib().mgbvh_raytrace_tlas_device(
            self.handle, _dv(tlas_nodes), _dv(tlas_perm), _dv(instances), ninstances, _dv(blas),
            nblas, _dv(blas_nodes), _dv(pos), vstride, _dv(tris_perm), _dv(blas_perm), _dv(rays),
            nrays, _dv(hits)))

    def wide8_collapse(self, nodes, lens, nnodes: int, wide, wide_cap: int, tri_index,
                       tri_cap: int, c_node: float = 1.0, c_prim: float = 0.3, p_max: int = 3):
        """Compressed 8-wide tree from device binary nodes.  Returns (nwide, cost)."""
        prm = Wide8Params(c_node, c_prim, p_max, 0)
        nw, cost = C.c_uint64(), C.c_float()
        check(lib().mgbvh_wide8_collapse_device(self.handle, _dv(nodes), _dv(lens), nnodes,
                                                C.byref(prm), _dv(wide), wide_cap,
                                                _dv(tri_index), tri_cap, C.byref(nw),
                                                C.byref(cost)))
        return nw.value, cost.value

    def wide8_collapse_batch(self, nodes, lens, nnodes: int, blas: np.ndarray,
                             node_counts: np.ndarray, wide, wide_cap: int, tri_index,
                             tri_cap: int, c_node: float = 1.0, c_prim: float = 0.3,
                             p_max: int = 3):
        """Every BLAS of a batched build collapsed at once.  blas: BLAS_DTYPE
        array, node_counts: u32 per BLAS (host).  Returns (first, count, total)."""
        prm = Wide8Params(c_node, c_prim, p_max, 0)
        blas = np.ascontiguousarray(blas)
        node_counts = np.ascontiguousarray(node_counts, dtype=np.uint32)
        nb = blas.shape[0]
        first = np.zeros(nb, np.uint32)
        count = np.zeros(nb, np.uint32)
        tot = C.c_uint64()
        check(lib().mgbvh_wide8_collapse_batch_device(
            self.handle, _dv(nodes), _dv(lens), nnodes, ptr(blas), ptr(node_counts), nb,
            C.byref(prm), _dv(wide), wide_cap, _dv(tri_index), tri_cap, ptr(first), ptr(count),
            C.byref(tot)))
        return first, count, tot.value

    def trace_wide8(self, wide, nwide: int, tri_index, pos, vstride, tris_perm, perm, rays,
                    nrays, hits):
        check(lib().mgbvh_trace_wide8_device(self.handle, _dv(wide), nwide, _dv(tri_index),
                                             _dv(pos), vstride, _dv(tris_perm), _dv(perm),
                                             _dv(rays), nrays, _dv(hits)))

    def gen_primary(self, cam: "Camera", width: int, height: int, rays):
        check(lib().mgbvh_gen_primary_rays_device(self.handle, C.byref(cam), width, height,
                                                  _dv(rays)))

    def gen_view(self, inv_view_proj, position, width: int, height: int, rays,
                 frame_count: int = 0, accumulate: bool = False):
        """The path tracer's primary rays from a column-major inverse
        view-projection matrix (16 floats) and the camera position."""
        m = np.ascontiguousarray(np.asarray(inv_view_proj, np.float32).reshape(16))
        p = np.ascontiguousarray(np.asarray(position, np.float32).reshape(3))
        check(lib().mgbvh_gen_view_rays_device(self.handle, ptr(m), ptr(p), width, height,
                                               frame_count, 1 if accumulate else 0, _dv(rays)))

    def gen_bounce(self, rays_in, hits, nrays, instances, ninstances, blas, nblas, pos, vstride,
                   tris_perm, seed, rays_out):
        check(lib().mgbvh_gen_bounce_rays_device(
            self.handle, _dv(rays_in), _dv(hits), nrays, _dv(instances), ninstances, _dv(blas),
            nblas, _dv(pos), vstride, _dv(tris_perm), seed, _dv(rays_out)))

    # ---- batched BLAS build (device buffers) ---------------------------------
    def build_blas_batch(self, pos, vstride, nverts, tris, meshes: np.ndarray, settings: Settings,
                         nodes, lens, perm, tris_out=None, splits=None, mesh_bounds=None):
        """meshes: MESH_DTYPE array (host).  Returns (blas BLAS_DTYPE array,
        node counts, stats dict)."""
        meshes = np.ascontiguousarray(meshes, dtype=MESH_DTYPE)
        nm = meshes.shape[0]
        blas = np.zeros(max(nm, 1), dtype=BLAS_DTYPE)
        counts = np.zeros(max(nm, 1), dtype=np.uint32)
        mb = None if mesh_bounds is None else np.ascontiguousarray(mesh_bounds, dtype=np.float32)
        st = Stats()
        check(lib().mgbvh_build_blas_batch_device(
            self.handle, _dv(pos), vstride, nverts, _dv(tris), ptr(meshes), nm, ptr(mb),
            C.byref(settings), _dv(nodes), _dv(lens), _dv(perm), _dv(tris_out), _dv(splits),
            ptr(blas), ptr(counts), C.byref(st)))
        return blas[:nm], counts[:nm], st.as_dict()

    def instance_boxes(self, blas_nodes, blas, nblas, object_to_world, instance_blas, ninst,
                       boxes):
        check(lib().mgbvh_instance_boxes_device(self.handle, _dv(blas_nodes), _dv(blas), nblas,
                                                _dv(object_to_world), _dv(instance_blas), ninst,
                                                _dv(boxes)))

    def build_aabbs_device(self, boxes, n, settings: Settings, nodes, lens, perm, splits=None):
        nn = C.c_uint64()
        check(lib().mgbvh_build_aabbs_device(self.handle, _dv(boxes), n, C.byref(settings),
                                             _dv(nodes), _dv(lens), _dv(perm), _dv(splits),
                                             C.byref(nn), None))
        return nn.value

    def build_tris_device(self, pos, vstride, nverts, tris, n, settings: Settings, nodes, lens,
                          perm, tris_out=None, splits=None, pos_offset=0, tri_offset=0,
                          node_offset=0):
        """Device build of one mesh whose arrays start at the given element
        offsets of larger device buffers.  Returns the node count."""
        nn = C.c_uint64()
        def at(b, off, size):
            return None if b is None else C.c_void_p(b.ptr.value + off * size)
        check(lib().mgbvh_build_tris_device(
            self.handle, at(pos, pos_offset, vstride), vstride, nverts, at(tris, tri_offset, 12),
            n, C.byref(settings), at(nodes, node_offset, 32), at(lens, node_offset, 4),
            at(perm, tri_offset, 4), at(tris_out, tri_offset, 12), at(splits, node_offset, 16),
            C.byref(nn), None))
        return nn.value

    def tree_sah(self, nodes, lens, blas: np.ndarray, node_counts: np.ndarray) -> np.ndarray:
        blas = np.ascontiguousarray(blas, dtype=BLAS_DTYPE)
        counts = np.ascontiguousarray(node_counts, dtype=np.uint32)
        out = np.zeros(max(len(blas), 1), dtype=np.float64)
        check(lib().mgbvh_tree_sah_device(self.handle, _dv(nodes), _dv(lens), ptr(blas),
                                          ptr(counts), len(blas), ptr(out)))
        return out[:len(blas)]

    def collect_samples(self, pos, vstride, nverts, tris, n, settings: Settings, min_objects,
                        feats, labels, max_samples) -> int:
        cnt = C.c_uint32()
        check(lib().mgbvh_collect_samples_device(self.handle, _dv(pos), vstride, nverts, _dv(tris),
                                                 n, C.byref(settings), min_objects, _dv(feats),
                                                 _dv(labels), max_samples, C.byref(cnt)))
        return cnt.value

    # ---- host-buffer builds -------------------------------------------------
    def build_tris(self, positions: np.ndarray, tris: np.ndarray, settings: Settings,
                   want_splits: bool = False, want_stats: bool = True):
        """Builds over host arrays.  `tris` (N,3) uint32 is permuted in place.
        Returns (nodes, lens, perm, splits, stats)."""
        assert positions.dtype == np.float32 and positions.flags["C_CONTIGUOUS"]
        assert tris.dtype == np.uint32 and tris.flags["C_CONTIGUOUS"]
        n = tris.shape[0]
        vstride = positions.strides[0] if positions.ndim == 2 else 12
        nslots = max(2 * n - 1, 1)
        nodes = np.zeros(nslots, dtype=NODE_DTYPE)
        lens = np.zeros(nslots, dtype=np.uint32)
        perm = np.zeros(max(n, 1), dtype=np.uint32)
        splits = np.zeros(nslots, dtype=SPLIT_DTYPE) if want_splits else None
        nn = C.c_uint64()
        st = Stats()
        check(lib().mgbvh_build_tris(self.handle, ptr(positions), vstride, positions.shape[0],
                                     ptr(tris), n, C.byref(settings), ptr(nodes), ptr(lens),
                                     C.byref(nn), ptr(perm), ptr(splits),
                                     C.byref(st) if want_stats else None))
        k = nn.value
        return (nodes[:k], lens[:k], perm[:n], None if splits is None else splits[:k],
                st.as_dict() if want_stats else None)

    def build_aabbs(self, boxes: np.ndarray, settings: Settings, want_splits: bool = False,
                    want_stats: bool = True):
        """Builds over (N,6) float32 boxes [min xyz, max xyz].
        Returns (nodes, lens, perm, splits, stats)."""
        boxes = np.ascontiguousarray(boxes, dtype=np.float32)
        n = boxes.shape[0]
        nslots = max(2 * n - 1, 1)
        nodes = np.zeros(nslots, dtype=NODE_DTYPE)
        lens = np.zeros(nslots, dtype=np.uint32)
        perm = np.zeros(max(n, 1), dtype=np.uint32)
        splits = np.zeros(nslots, dtype=SPLIT_DTYPE) if want_splits else None
        nn = C.c_uint64()
        st = Stats()
        check(lib().mgbvh_build_aabbs(self.handle, ptr(boxes), n, C.byref(settings), ptr(nodes),
                                      ptr(lens), C.byref(nn), ptr(perm), ptr(splits),
                                      C.byref(st) if want_stats else None))
        k = nn.value
        return (nodes[:k], lens[:k], perm[:n], None if splits is None else splits[:k],
                st.as_dict() if want_stats else None)
