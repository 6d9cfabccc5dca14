"""Builds tests/c/fg_conformance (gcc, strict C99) against include/focus_gpu.h,
libfocus_gpu.so and the oracle's liboracle.so.  Test infrastructure only."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(HERE, "fg_conformance")


def build(out: str = EXE, verbose: bool = False) -> str:
    pkg, orc = os.path.join(ROOT, "focus-stacker_amd"), os.path.join(ROOT, "oracle")
    rel = lambda d: os.path.relpath(d, os.path.dirname(os.path.abspath(out)))  # noqa: E731
    cmd = ["gcc", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O1", "-D_POSIX_C_SOURCE=200809L",
           "-I", os.path.join(ROOT, "include"), "-I", orc, os.path.join(HERE, "fg_conformance.c"),
           "-L", pkg, "-L", orc, "-lfocus_gpu", "-loracle", "-lpthread", "-lm",
           f"-Wl,-rpath,$ORIGIN/{rel(pkg)}:$ORIGIN/{rel(orc)}", "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
