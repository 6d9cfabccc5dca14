"""The ABI from C: tests/c/fg_conformance.c, built with gcc as strict C99 against
include/focus_gpu.h and libfocus_gpu.so, makes the calls the Zig glue makes (the
reference's Zig binds C through @cImport, which sees what a C compiler sees).

CPU: struct layouts and constants (compile-time), null-argument and thread-local error
paths, the overlap ROI against the oracle, the TIFF writer.  GPU: the pass-1/pass-2 loop
of every method, the uncached second loop, the one-shot call, aligned frames, the size /
format / order errors, remap + aligned TIFF and the server, each against the oracle."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "c"))


def test_conformance_host(tmp_path):
    import build as cbuild
    exe = cbuild.build(out=str(tmp_path / "fg_conformance"))
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout


@pytest.mark.gpu
def test_conformance_gpu(tmp_path):
    import build as cbuild
    if not os.path.exists(cbuild.EXE):
        pytest.fail(f"{cbuild.EXE} is not built: run __graft_entry__.build()")
    print("running fg_conformance --gpu", flush=True)
    r = subprocess.run([cbuild.EXE, "--gpu", str(tmp_path)], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed (with GPU)" in r.stdout
