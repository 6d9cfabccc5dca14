"""Builds libfocus_gpu.so (HIP for gfx950) in-tree with hipcc.

Flags that matter for results:
  -ffp-contract=off            no FMA contraction: the reference's float mode is strict,
                               and every f32 stage must round where the reference rounds
  -target-feature -packed-fp32-ops   keep the compiler from emitting packed FP32 math
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libfocus_gpu.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in (
    "fg_kernels.hip", "fg_pass.hip", "fg_single.hip", "fg_noalloc.hip", "fg_api.cpp", "fg_server.cpp",
    "fg_batch.cpp", "fg_tiff.cpp")]
FLAGS = [
    "-Wl,--no-undefined",
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-x", "hip",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
    "-I", os.path.join(ROOT, "include"),
]


def build(verbose: bool = False) -> str:
    cmd = ["hipcc", *FLAGS, *SOURCES, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
