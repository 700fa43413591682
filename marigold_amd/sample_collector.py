"""BVH profiling-sample collector, command-line compatible with marigold's
bvh_sample_collector (bvh_sample_collector/src/main.rs:9-70):

    python -m marigold_amd.sample_collector <mesh_dir> <output_dir> <generation_name> <num_samples>

Every entry of <mesh_dir> is a glTF scene directory.  Each of its meshes is
built <num_samples> times with the mesh's own bounds, max_depth 24 and
profiling on, so every build writes bvh_<scene>_<mesh index>_<id>.json into
<output_dir>/<generation_name> (the JSON of bvh/src/lib.rs:836-866, written
by the library).  The builds run on the GPU through mgbvh_build_tris, one
after another on one context (the reference spreads them over rayon threads;
each build is single-threaded there).  Mesh indices follow the loader's
(mesh, primitive) order.
"""
from __future__ import annotations

import os
import sys

from . import _native as N
from . import gltf

BLAS_SAMPLE_DEPTH = 24  # main.rs:50


def collect(mesh_dir: str, output_dir: str, generation_name: str, num_samples: int,
            ctx: N.Context | None = None, out=print) -> int:
    """Runs the collection; returns the number of builds."""
    own = ctx is None
    ctx = ctx or N.Context(0)
    builds = 0
    try:
        for entry in sorted(os.listdir(mesh_dir)):
            out(f"loading gltf scene {entry}...")
            meshes, _ = gltf.load(os.path.join(mesh_dir, entry))
            out(f"finished loading gltf scene {entry}!")
            jdir = os.path.join(output_dir, generation_name)
            for sample in range(num_samples):
                for index, m in enumerate(meshes):
                    s = N.default_settings(max_depth=BLAS_SAMPLE_DEPTH, bounds=m.bounds,
                                           name=f"{entry}_{index}", profiling_info=True,
                                           profiling_info_directory=jdir)
                    tris = m.triangles.copy()
                    ctx.build_tris(m.vertices, tris, s, want_stats=True)
                    builds += 1
                out(f"\tsample {sample + 1} done for {entry}!")
    finally:
        if own:
            ctx.close()
    return builds


def main(argv=None) -> int:
    args = list(sys.argv[1:] if argv is None else argv)
    if len(args) != 4:
        print(__doc__.strip().splitlines()[3].strip(), file=sys.stderr)
        return 2
    collect(args[0], args[1], args[2], int(args[3]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
