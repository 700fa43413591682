"""C ABI checks that need no GPU: the built library loads, exports every entry
point include/mgbvh.h declares, agrees with the ctypes layouts, and fails
cleanly (no abort) when no gfx950 device is present."""
import ctypes as C
import os
import re
import subprocess

import pytest

from marigold_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mgbvh.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mgbvh_[a-z0-9_]+)\s*\(", src)))


def test_header_functions_exported():
    lib = N.lib()
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (mgbvh_[a-z0-9_]+)", out))
    assert set(names) <= exported


def test_struct_layouts_match_header(tmp_path):
    prog = tmp_path / "layout.c"
    prog.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "mgbvh.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(mgbvh_node32), sizeof(mgbvh_aabb), sizeof(mgbvh_split_rec),
         sizeof(mgbvh_settings), sizeof(mgbvh_stats));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(mgbvh_settings, max_depth), offsetof(mgbvh_settings, method),
         offsetof(mgbvh_settings, profiling_info_directory), offsetof(mgbvh_stats, height),
         offsetof(mgbvh_stats, n_guard_fallbacks), offsetof(mgbvh_node32, child_node));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), str(prog), "-o", str(exe)],
                   check=True)
    a, b = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")[:2]
    sizes = [int(x) for x in a.split()]
    offs = [int(x) for x in b.split()]
    assert sizes == [32, 24, 16, C.sizeof(N.Settings), C.sizeof(N.Stats)]
    assert offs == [N.Settings.max_depth.offset, N.Settings.method.offset,
                    N.Settings.profiling_info_directory.offset, N.Stats.height.offset,
                    N.Stats.n_guard_fallbacks.offset, 28]


def test_settings_defaults():
    s = N.Settings()
    N.lib().mgbvh_settings_default(C.byref(s))
    assert (s.max_depth, s.min_leaf_objects, s.max_leaf_objects, s.method, s.has_bounds) == (32, 1, 1, 0, 0)


def test_null_arguments_fail_cleanly():
    L = N.lib()
    assert L.mgbvh_ctx_create(0, None) == N.ERR_INVALID_ARG
    assert L.mgbvh_ctx_synchronize(None) == N.ERR_INVALID_ARG
    assert L.mgbvh_build_tris(None, None, 12, 0, None, 0, None, None, None, None, None, None, None) \
        == N.ERR_INVALID_ARG
    assert "NULL" in N.last_error()


def _has_gpu():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_has_gpu(), reason="a GPU is present; covered by the gpu tests")
def test_no_device_error():
    with pytest.raises(N.MgbvhError) as e:
        N.Context(0)
    assert e.value.code == N.ERR_NO_DEVICE


def test_trace_struct_layouts_match_header(tmp_path):
    prog = tmp_path / "layout2.c"
    prog.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "mgbvh.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(mgbvh_ray), sizeof(mgbvh_hit), sizeof(mgbvh_blas_desc),
         sizeof(mgbvh_instance), sizeof(mgbvh_camera));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(mgbvh_ray, dir), offsetof(mgbvh_ray, tmax),
         offsetof(mgbvh_hit, tri_slot), offsetof(mgbvh_blas_desc, vertex_offset),
         offsetof(mgbvh_instance, blas), offsetof(mgbvh_camera, aspect));
  return 0;
}
''')
    exe = tmp_path / "layout2"
    subprocess.run(["gcc", "-std=c11", "-I" + os.path.join(ROOT, "include"), str(prog), "-o", str(exe)],
                   check=True)
    a, b = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")[:2]
    assert [int(x) for x in a.split()] == [N.RAY_DTYPE.itemsize, N.HIT_DTYPE.itemsize,
                                           N.BLAS_DTYPE.itemsize, N.INSTANCE_DTYPE.itemsize,
                                           C.sizeof(N.Camera)]
    assert [int(x) for x in b.split()] == [N.RAY_DTYPE.fields["dir"][1], N.RAY_DTYPE.fields["tmax"][1],
                                           N.HIT_DTYPE.fields["tri_slot"][1],
                                           N.BLAS_DTYPE.fields["vertex_offset"][1],
                                           N.INSTANCE_DTYPE.fields["blas"][1], N.Camera.aspect.offset]


def test_trace_null_arguments_fail_cleanly():
    L = N.lib()
    assert L.mgbvh_trace_closest_device(None, None, None, None, 12, None, None, None, 1, None) \
        == N.ERR_INVALID_ARG
    assert L.mgbvh_trace_scene_device(None, None, None, None, None, 0, None, 0, None, None, None,
                                      12, None, None, None, 1, None) == N.ERR_INVALID_ARG
    assert L.mgbvh_gen_primary_rays_device(None, None, 4, 4, None) == N.ERR_INVALID_ARG
