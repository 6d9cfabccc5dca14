"""GPU: fg_fuse_stack_device_noalloc (working buffers in the library's fixed scratch area, no device
allocation) through the C ABI, against fg_fuse_stack, the one-shot entry point it varies: the same
frames, poses and options must give the same output bits.  Two calls in a row, the second on a
smaller stack (the area keeps the first call's contents), a stack whose frame cache does not fit
the area (the uncached second pass), and a stack too large for the area (a clear error)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fg():
    import focus_gpu
    return focus_gpu


@pytest.fixture(scope="module")
def ctx(fg):
    c = fg.Context(0)
    yield c
    c.close()


def _stack(config_id, frames, width, height, device="cpu", **kw):
    import synth
    spec = synth.config_spec(config_id, frames=frames, width=width, height=height, **kw)
    st = synth.Stack(spec, device=device)
    return [st.frame_numpy(k) for k in range(frames)], st.poses


def _reference(fg, ctx, frames, poses, roi, method, sample_type):
    """fg_fuse_stack on host frames: the existing one-shot entry point."""
    n = len(frames)
    descs = (fg.ImageDesc * n)(*[fg.desc_for(f) for f in frames])
    ptrs = (C.c_void_p * n)(*[f.ctypes.data for f in frames])
    pz = (fg.Pose * n)(*[fg.pose(p) for p in poses])
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    o = fg.StackOpts(method, 5, 0.35, n, w, h, sample_type, 0, -1.0, -1, None)
    out = np.zeros((h, w, 3), frames[0].dtype)
    fg.check(fg.lib().fg_fuse_stack(ctx.h, n, C.cast(descs, C.c_void_p), C.cast(ptrs, C.c_void_p),
                                    C.cast(pz, C.c_void_p), C.byref(fg.rect(roi)), C.byref(o), fg._p(out)))
    return out


def _noalloc(fg, ctx, frames, poses, roi, method):
    d_frames = [ctx.upload(f) for f in frames]
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    d_out = ctx.alloc(w * h * 3 * frames[0].itemsize)
    try:
        ctx.fuse_stack_device_noalloc(d_frames, [fg.desc_for(f) for f in frames], [fg.pose(p) for p in poses], roi,
                                      d_out, method=method)
        return ctx.download(d_out, (h, w, 3), frames[0].dtype)
    finally:
        for d in d_frames:
            ctx.free(d)
        ctx.free(d_out)


def test_noalloc_equals_fuse_stack_on_two_calls(ctx, fg):
    cases = [("hybrid u16, 4 x 300x200", dict(frames=4, width=300, height=200, dtype="u16"), fg.METHOD_HYBRID,
              fg.SAMPLE_U16),
             ("pyramid u8, 3 x 150x110", dict(frames=3, width=150, height=110), fg.METHOD_PYRAMID, fg.SAMPLE_U8)]
    for name, kw, method, st in cases:
        print(f"noalloc call: {name}", flush=True)
        frames, poses = _stack(0, **kw)
        roi = fg.common_overlap_roi([(kw["width"], kw["height"])] * kw["frames"], [fg.pose(p) for p in poses])
        got = _noalloc(fg, ctx, frames, poses, roi, method)
        ref = _reference(fg, ctx, frames, poses, roi, method, st)
        assert np.array_equal(got, ref), f"{name}: {int((got != ref).sum())} samples differ"


def test_noalloc_uncached_in_the_area_equals_fuse_stack(ctx, fg):
    """Ten 1600x1200 frames: their cache (about 25 MB a frame) exceeds the area, so the stack runs
    the uncached second pass, fed again from the device frames."""
    import torch
    print("noalloc call: pyramid u8, 10 x 1600x1200 (uncached)", flush=True)
    frames, poses = _stack(0, 10, 1600, 1200, device=torch.device("cuda:0"), shift=5.0)
    roi = fg.common_overlap_roi([(1600, 1200)] * 10, [fg.pose(p) for p in poses])
    got = _noalloc(fg, ctx, frames, poses, roi, fg.METHOD_PYRAMID)
    ref = _reference(fg, ctx, frames, poses, roi, fg.METHOD_PYRAMID, fg.SAMPLE_U8)
    assert np.array_equal(got, ref)


def test_noalloc_refuses_a_stack_beyond_the_area(ctx, fg):
    n, w, h = 2, 4000, 3000
    d = ctx.alloc(w * h * 3)
    try:
        desc = fg.ImageDesc(w, h, 3, 0, fg.SAMPLE_U8)
        with pytest.raises(fg.FgError) as e:
            ctx.fuse_stack_device_noalloc([d] * n, [desc] * n, None, None, d, method=fg.METHOD_PYRAMID)
        assert e.value.code == -1 and "scratch area holds" in str(e.value)
    finally:
        ctx.free(d)
