"""GPU: whole stacks at the real frame size, 7952x5304, through the C ABI.

The oracle takes minutes per frame at this size, so these tests check properties that fix the
output whatever the size:
- N identical frames reproduce the input. Every mask level is then 1/N, and the collapse of a
  Laplacian pyramid returns its base (src/fuse/pyramid.zig:223-269, 303-316).
- The hybrid at sharpness 0 returns the pyramid image. The merge then keeps the soft Laplacian
  (mergeHybridImages src/fuse/pyramid.zig:271-301), and the soft image is the pyramid result.
Both hold to well under half a sample step, so the quantized outputs must match exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 7952, 5304


@pytest.fixture(scope="module")
def fg():
    import focus_gpu
    return focus_gpu


@pytest.fixture(scope="module")
def ctx(fg):
    c = fg.Context(0)
    yield c
    c.close()


def _frame(dtype, seed):
    """Smooth random texture (so that the contrast weights vary) plus fine noise."""
    rng = np.random.default_rng(seed)
    hi = 255 if dtype == np.uint8 else 65535
    coarse = rng.integers(0, int(0.8 * hi) + 1, (H // 64 + 1, W // 64 + 1, 3), dtype=dtype)
    base = np.repeat(np.repeat(coarse, 64, axis=0), 64, axis=1)[:H, :W]
    return np.ascontiguousarray(base + rng.integers(0, int(0.2 * hi) + 1, (H, W, 3), dtype=dtype))


@pytest.mark.parametrize("dtype,posed,method", [(np.uint8, False, "pyramid"), (np.uint8, True, "pyramid"),
                                                (np.uint16, False, "pyramid"), (np.uint8, False, "hybrid"),
                                                (np.uint8, True, "hardmask"), (np.uint16, False, "softmask")])
def test_identical_frames_reproduce_the_input(ctx, fg, dtype, posed, method):
    """Four copies of one frame: aligned input, or remapped under the identity pose (every sample
    on a pixel centre, so the remap returns the frame) over the full frame.  Hardmask keeps the
    first frame on equal weights; softmask's weighted mean of equal samples is the sample."""
    f = _frame(dtype, 7)
    st = fg.SAMPLE_U8 if dtype == np.uint8 else fg.SAMPLE_U16
    m = {"pyramid": fg.METHOD_PYRAMID, "hybrid": fg.METHOD_HYBRID, "hardmask": fg.METHOD_HARDMASK,
         "softmask": fg.METHOD_SOFTMASK}[method]
    print(f"full size identical frames {np.dtype(dtype).name} posed={posed} {method}", flush=True)
    s = fg.Stack(ctx, 4, W, H, st, method=m)
    for _ in range(4):
        if posed:
            s.add_frame(f, fg.pose({}), (0, 0, W, H))
        else:
            s.add_frame(f)
    out = s.finish()
    if method in ("pyramid", "hybrid"):
        assert len(s.levels()) == 14
    s.close()
    assert np.array_equal(out, f)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_hybrid_at_sharpness_zero_is_the_pyramid(ctx, fg, dtype):
    """Three different frames under small similarity poses (the homography remap path), fused by
    pyramid contrast and by the hybrid at sharpness 0."""
    frames = [_frame(dtype, 11 + k) for k in range(3)]
    st = fg.SAMPLE_U8 if dtype == np.uint8 else fg.SAMPLE_U16
    poses = [fg.pose({}), fg.pose(dict(yaw=0.0004, roll=0.001, hfov_delta=0.02)),
             fg.pose(dict(pitch=-0.0003, roll=-0.0008, hfov_delta=-0.015))]
    roi = fg.common_overlap_roi([(W, H)] * 3, poses)
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    print(f"full size hybrid vs pyramid {np.dtype(dtype).name} roi={roi}", flush=True)
    outs = {}
    for name, m in (("pyramid", fg.METHOD_PYRAMID), ("hybrid", fg.METHOD_HYBRID)):
        s = fg.Stack(ctx, 3, w, h, st, method=m, sharpness=0.0)
        for f, p in zip(frames, poses):
            s.add_frame(f, p, roi)
        outs[name] = s.finish()
        s.close()
    assert np.array_equal(outs["hybrid"], outs["pyramid"])
