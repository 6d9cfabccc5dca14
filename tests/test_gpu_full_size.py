"""GPU: whole stacks at the real frame size, 7952x5304, through the C ABI.

Against the oracle (its loops run on every host thread the job has, with the same bits as one
thread): short stacks from the seeded generators of the 30-frame u8 and the 143-frame u16
workloads (the same poses: shifts of +-20 px, roll +-0.3 deg, scale +-1 %; defocus up to 6 px;
the u16 one with its three focal planes), at full pyramid depth and capped at 9 levels, pyramid
and hybrid (sharpness 0.35), with the intermediate sums read back.  The frames are generated on
the GPU and copied to the host for the oracle.

Properties that fix the output whatever the size:
- N identical frames reproduce the input. Every mask level is then 1/N, and the collapse of a
  Laplacian pyramid returns its base (src/fuse/pyramid.zig:223-269, 303-316).
- The hybrid at sharpness 0 returns the pyramid image. The merge then keeps the soft Laplacian
  (mergeHybridImages src/fuse/pyramid.zig:271-301), and the soft image is the pyramid result.
Both hold to well under half a sample step, so the quantized outputs must match exactly.
"""
import os

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


# --------------------------------------------------------------------------
# against the oracle
# --------------------------------------------------------------------------

def _img_close(a, b, max_abs=1, mean_abs=0.05):
    d = np.abs(a.astype(np.int64) - b.astype(np.int64))
    assert d.max() <= max_abs, f"max |d| = {d.max()}"
    assert d.mean() < mean_abs, f"mean |d| = {d.mean()}"
    return d


def _level_close(got, ref, rel=1e-5):
    rms = float(np.sqrt(np.mean(np.square(ref.astype(np.float64)))))
    err = np.abs(got.astype(np.float64) - ref.astype(np.float64))
    bad = err > rel * np.maximum(np.abs(ref.astype(np.float64)), rms) + 1e-30
    assert not bad.any(), f"{bad.sum()} of {bad.size} values off; worst {err.max()} (rms {rms})"


def _generated(config_id, frames):
    """Frames 0..frames-1 of a stack of `frames` from the workload's seeded generator, made on the
    GPU, as host arrays; the poses; the common-overlap ROI."""
    import torch
    import synth
    import focus_gpu
    spec = synth.config_spec(config_id, frames=frames)
    st = synth.Stack(spec, device=torch.device("cuda:0"))
    out = [st.frame_numpy(k) for k in range(frames)]
    poses = st.poses
    del st
    torch.cuda.empty_cache()
    roi = focus_gpu.common_overlap_roi([(spec.width, spec.height)] * frames, [focus_gpu.pose(p) for p in poses])
    return out, poses, roi


@pytest.fixture(scope="module")
def gen_u8():
    print("generating 4 frames of the 7952x5304 u8 workload on the GPU", flush=True)
    return _generated(1, 4)


@pytest.fixture(scope="module")
def gen_u16():
    print("generating 3 frames of the 7952x5304 u16 three-plane workload on the GPU", flush=True)
    return _generated(2, 3)


def _oracle():
    from oracle import Oracle
    o = Oracle()
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    o.set_threads(n)
    return o


def _gpu_stack(fg, ctx, frames, poses, roi, **kw):
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    st = fg.SAMPLE_U8 if frames[0].dtype == np.uint8 else fg.SAMPLE_U16
    s = fg.Stack(ctx, len(frames), w, h, st, **kw)
    for f, p in zip(frames, poses):
        s.add_frame(f, fg.pose(p), roi)
    assert not s.needs_pass2_frames()
    return s, s.finish()


@pytest.mark.parametrize("levels", [0, 9])
def test_u8_workload_frames_match_oracle(ctx, fg, gen_u8, levels):
    """Four frames of the 30 x 7952x5304 u8 workload's generator, remap + pyramid-contrast at the
    reference's depth (14 levels) and at the 9 levels the workload names."""
    from oracle import make_pose
    frames, poses, roi = gen_u8
    print(f"u8 workload, 4 frames, roi={roi}, levels={levels}", flush=True)
    s, out = _gpu_stack(fg, ctx, frames, poses, roi, max_levels=levels)
    assert len(s.levels()) == (14 if levels == 0 else 9)
    o = _oracle()
    ref, dbg = o.stack(frames, [make_pose(**p) for p in poses], roi, max_levels=levels, debug=True)
    d = _img_close(out, ref)
    print(f"  max {d.max()} mean {d.mean():.2e} differing {float((d != 0).mean()):.2e}", flush=True)
    assert np.array_equal(s.readback("union_support"), dbg["union_support"])
    _level_close(s.readback("norm_sum"), dbg["norm_sum"])
    for k in (0, 3):
        _level_close(s.readback("weights", k), dbg["weights"][k])
    s.close()


@pytest.mark.parametrize("method", ["pyramid", "hybrid"])
def test_u16_workload_frames_match_oracle(ctx, fg, gen_u16, method):
    """Three frames of the 143 x 7952x5304 u16 three-plane workload's generator: pyramid-contrast,
    and hybrid-pyramid-contrast at sharpness 0.35 (its soft and hard images checked too)."""
    from oracle import make_pose, METHOD_PYRAMID, METHOD_HYBRID
    frames, poses, roi = gen_u16
    print(f"u16 workload, 3 frames, roi={roi}, {method}", flush=True)
    m = fg.METHOD_PYRAMID if method == "pyramid" else fg.METHOD_HYBRID
    s, out = _gpu_stack(fg, ctx, frames, poses, roi, method=m, sharpness=0.35)
    o = _oracle()
    om = METHOD_PYRAMID if method == "pyramid" else METHOD_HYBRID
    ref, dbg = o.stack(frames, [make_pose(**p) for p in poses], roi, method=om, sharpness=0.35, debug=True)
    d = _img_close(out, ref)
    print(f"  max {d.max()} mean {d.mean():.2e} differing {float((d != 0).mean()):.2e}", flush=True)
    _level_close(s.readback("norm_sum"), dbg["norm_sum"])
    if method == "hybrid":
        _img_close(s.readback("soft"), dbg["soft"])
        hard = s.readback("hard")
        differ = (hard != dbg["hard"]).any(axis=2)
        top2 = np.sort(dbg["weights"].astype(np.float64), axis=0)[-2:]
        tied = (top2[1] - top2[0]) <= 1e-6 * np.maximum(top2[1], 1e-30)
        print(f"  hard image: {int(differ.sum())} winners differ, {int((differ & ~tied).sum())} away from a tie",
              flush=True)
        assert not (differ & ~tied).any()
    s.close()


def test_u8_workload_hardmask_softmask_match_oracle(ctx, fg, gen_u8):
    """The single-pass methods on the same full-size frames: winners bit-exact, softmask within
    the image tolerance."""
    from oracle import make_pose, METHOD_HARDMASK, METHOD_SOFTMASK
    frames, poses, roi = gen_u8
    o = _oracle()
    opo = [make_pose(**p) for p in poses]
    for name, m, om in (("hardmask", fg.METHOD_HARDMASK, METHOD_HARDMASK),
                        ("softmask", fg.METHOD_SOFTMASK, METHOD_SOFTMASK)):
        print(f"u8 workload, 4 frames, {name}", flush=True)
        s, out = _gpu_stack(fg, ctx, frames, poses, roi, method=m)
        s.close()
        ref = o.stack(frames, opo, roi, method=om)
        if name == "hardmask":
            assert np.array_equal(out, ref), f"{int((out != ref).any(axis=2).sum())} winners differ"
        else:
            _img_close(out, ref)


# --------------------------------------------------------------------------
# the whole 143-frame u16 workload
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def gen_u16_full(fg):
    """All 143 frames of the u16 three-plane workload, 7952x5304, left on the GPU (36 GB)."""
    import torch
    import synth
    spec = synth.config_spec(2)
    print(f"generating {spec.frames} frames of the 7952x5304 u16 workload on the GPU", flush=True)
    st = synth.Stack(spec, device=torch.device("cuda:0"))
    frames = []
    for k in range(spec.frames):
        f = st.frame(k)
        frames.append((f - (f > 32767).to(torch.int32) * 65536).to(torch.int16).contiguous())
    poses = [fg.pose(p) for p in st.poses]
    del st
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    roi = fg.common_overlap_roi([(spec.width, spec.height)] * spec.frames, poses)
    yield frames, poses, roi
    del frames
    torch.cuda.empty_cache()


@pytest.mark.parametrize("method", ["pyramid", "hybrid"])
def test_u16_workload_143_frames_cached_equals_uncached(ctx, fg, gen_u16_full, method):
    """The whole 143-frame stack, streamed: with every frame cached (the benchmark's form: the
    hybrid keeps its winners as frame slots 0..142) and under a budget too small for the frame cache
    (the reference's uncached second loop, every frame fed twice; the hybrid keeps winner colours).
    Both write the same image, and the norm sum and union support agree."""
    import torch
    frames, poses, roi = gen_u16_full
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    m = fg.METHOD_PYRAMID if method == "pyramid" else fg.METHOD_HYBRID
    desc = fg.ImageDesc(W, H, 3, 0, fg.SAMPLE_U16)
    outs, sums = {}, {}
    for name, frac in (("cached", 0.8), ("uncached", 0.02)):
        print(f"143 x u16 {method} {name}, roi={roi}", flush=True)
        s = fg.Stack(ctx, len(frames), w, h, fg.SAMPLE_U16, method=m, sharpness=0.35, memory_fraction=frac)
        for f, p in zip(frames, poses):
            s.add_frame_device(f.data_ptr(), desc, p, roi)
        assert s.needs_pass2_frames() == (name == "uncached")
        if s.needs_pass2_frames():
            for f, p in zip(frames, poses):
                s.pass2_frame_device(f.data_ptr(), desc, p, roi)
        out = torch.empty((h, w, 3), dtype=torch.int16, device="cuda:0")
        torch.cuda.synchronize()
        s.finish_device(out.data_ptr())
        s.sync()
        outs[name] = out.cpu().numpy()
        sums[name] = (s.readback("norm_sum"), s.readback("union_support"))
        s.close()
    assert np.array_equal(outs["cached"], outs["uncached"])
    assert np.array_equal(sums["cached"][0], sums["uncached"][0])
    assert np.array_equal(sums["cached"][1], sums["uncached"][1])
    assert int(outs["cached"].view(np.uint16).max()) > 0
