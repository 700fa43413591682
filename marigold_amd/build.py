"""Build recipe of libmgbvh.so (hipcc, gfx950, in-tree).

Flags that matter for bit-exactness with the reference builder:
  -ffp-contract=off     no FMA contraction of SAH cost expressions
  (no -ffast-math)      IEEE division/compare semantics, correctly rounded fp32 division
Packed-FP32 instruction selection is disabled (-packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmgbvh.so")

SOURCES = ["mgbvh_build.hip", "mgbvh_api.hip", "mgbvh_trace.hip", "mgbvh_scene.hip", "mgbvh_learned.hip",
           "mgbvh_wide8.hip", "mgbvh_noalloc.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    return "hipcc"


def build_native(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmgbvh.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_native(verbose=True))
