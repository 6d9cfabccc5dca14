"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/focus_oracle.c twice with gcc, as plain C with FMA contraction
disabled (the reference's float mode is strict) and OpenMP for its independent
per-pixel / per-row loops (same bits at any thread count):
  oracle/liboracle.so    OR_REAL=float   the reference's own f32/f64 mix
  oracle/liboracle64.so  OR_REAL=double  f32 storage promoted to f64
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FLAGS = ["-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
         "-Wall", "-Wno-unused-function"]


def build(verbose: bool = False) -> None:
    src = os.path.join(HERE, "focus_oracle.c")
    for name, real in (("liboracle.so", "float"), ("liboracle64.so", "double")):
        out = os.path.join(HERE, name)
        cmd = ["gcc", *FLAGS, f"-DOR_REAL={real}", src, "-o", out + ".tmp", "-lm"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
