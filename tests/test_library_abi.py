"""CPU-side checks of the C ABI: the HIP library loads and exports every function that
include/focus_gpu.h declares, and the ctypes structs match the header's layout.
No compute call is made (there is no GPU here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "focus_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    fns = declared_functions()
    for f in ("fg_remap", "fg_stack_create", "fg_stack_add_frame", "fg_stack_finish", "fg_fuse_stack",
              "fg_common_overlap_roi", "fg_last_error"):
        assert f in fns


def test_library_exports_every_declared_symbol():
    import focus_gpu
    if not os.path.exists(focus_gpu.LIB_PATH):
        pytest.fail("libfocus_gpu.so not built (run __graft_entry__.build())")
    lib = C.CDLL(focus_gpu.LIB_PATH)
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_struct_layouts():
    import focus_gpu as f
    assert C.sizeof(f.Pose) == 15 * 8
    assert C.sizeof(f.Rect) == 16
    assert C.sizeof(f.ImageDesc) == 20
    assert C.sizeof(f.ContextOpts) == 8
    assert f.StackOpts.stream.offset == 40 and C.sizeof(f.StackOpts) == 48


def test_version_and_error_strings_without_gpu():
    import focus_gpu as f
    lib = f.lib()
    assert b"gfx950" in lib.fg_version()
    # the message is per thread: a fresh thread has none, whatever failed before on others
    import threading
    seen = []
    t = threading.Thread(target=lambda: seen.append(lib.fg_last_error()))
    t.start()
    t.join()
    assert seen == [b""]


def test_invalid_arguments_fail_before_touching_the_gpu():
    import focus_gpu as f
    lib = f.lib()
    assert lib.fg_context_create(None, None) == -1
    assert lib.fg_stack_create(None, None, None) == -1
    assert b"null" in lib.fg_last_error()
    out = f.Rect()
    assert lib.fg_common_overlap_roi(0, None, None, None, None, C.byref(out)) == -1


def test_common_overlap_roi_matches_oracle(oracle):
    """fg_common_overlap_roi is host arithmetic (no GPU): same ROI as the oracle on every pose path."""
    import numpy as np
    import focus_gpu as f
    from oracle import make_pose
    rng = np.random.default_rng(21)
    for trial in range(30):
        n = int(rng.integers(1, 8))
        w, h = int(rng.integers(20, 900)), int(rng.integers(20, 700))
        kind = trial % 3
        poses = []
        for k in range(n):
            kw = dict(yaw=rng.uniform(-0.02, 0.02), pitch=rng.uniform(-0.02, 0.02), roll=rng.uniform(-0.01, 0.01),
                      hfov_delta=rng.uniform(-1, 1))
            if kind == 1:
                kw.update(radial_a=rng.uniform(-1e-3, 1e-3), center_shift_x=rng.uniform(-1, 1))
            if kind == 2:
                kw.update(trans_x=rng.uniform(-0.1, 0.1), trans_z=rng.uniform(-0.01, 0.01))
            poses.append(kw if k else {})
        got = f.common_overlap_roi([(w, h)] * n, [f.pose(p) for p in poses])
        ref = oracle.common_overlap_roi([(w, h)] * n, [make_pose(**p) for p in poses])
        assert got == ref
