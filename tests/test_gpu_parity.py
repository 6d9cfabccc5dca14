"""GPU parity: the HIP path through the C ABI against the CPU oracle on the same seeded inputs.

Tolerances (stated by the task for this path):
  * 8/16-bit final images: max |d| <= 1 LSB and mean |d| < 0.05 LSB;
  * float pyramid / accumulator levels: |d| <= 1e-5 * max(|ref|, RMS of the level);
  * remap output and hard-mask winners: integer and index results, asserted bit-exact.  The GPU
    evaluates the reference's f64 remap formulas (the legacy trig chain as its equivalent
    homography) and picks winners with the same strict > over weights that round identically.
    The one exception: at 16 bits the contrast's f64 sum of squares is not exact, and the GPU adds
    the window in another order than the reference's sliding sum, so a weight may differ in its
    last bit; a winner may then differ only at a pixel where two frames' weights tie to within
    that rounding (_winners_exact_or_tied).
"""
import json
import os
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def fg():
    import focus_gpu
    return focus_gpu


@pytest.fixture(scope="module")
def ctx(fg):
    c = fg.Context(0)
    yield c
    c.close()


def _img_close(a, b, max_abs=1, mean_abs=0.05):
    d = np.abs(a.astype(np.int64) - b.astype(np.int64))
    assert d.max() <= max_abs, f"max |d| = {d.max()}"
    assert d.mean() < mean_abs, f"mean |d| = {d.mean()}"
    return d


def _level_close(got, ref, rel=1e-5):
    rms = float(np.sqrt(np.mean(np.square(ref.astype(np.float64))))) if ref.size else 0.0
    scale = np.maximum(np.abs(ref.astype(np.float64)), rms)
    err = np.abs(got.astype(np.float64) - ref.astype(np.float64))
    bad = err > rel * scale + 1e-30
    assert not bad.any(), f"{bad.sum()} of {bad.size} values off; worst {err.max()} (rms {rms})"


# --------------------------------------------------------------------------
# remap
# --------------------------------------------------------------------------

def _remap_close(got, ref):
    """Bit-exact, alpha included; the colour of a sample is compared where either side has
    support.  A source coordinate that lands exactly on the -1 boundary row/column (identity-like
    poses with the ROI outside the frame) comes out of the reference's trig chain as -1 +- 1 ulp,
    so there the colour of a zero-support sample is not defined by the reference either."""
    unsupported = (got[..., -1] == 0) & (ref[..., -1] == 0)
    g = got.astype(np.int64)
    r = ref.astype(np.int64)
    g[unsupported, :-1] = 0
    r[unsupported, :-1] = 0
    d = np.abs(g - r)
    assert not d.any(), f"{int((d != 0).sum())} samples differ, max {d.max()}"
    return d


def _winners_exact_or_tied(out, ref, weights, sample_bytes):
    """Hard-mask winners (updateRange src/fuse/blend.zig:237-288: the first frame with a strictly
    greater weight).  8-bit: every weight is exact in both, so the winners are identical.  16-bit:
    a pixel may differ only where the oracle's two largest weights are within 1e-6 of each other
    (the last-bit rounding of the f64 sum of squares)."""
    differ = (out != ref).any(axis=2)
    if sample_bytes == 1 or not differ.any():
        assert not differ.any(), f"{int(differ.sum())} winners differ"
        return
    top2 = np.sort(np.asarray(weights, np.float64), axis=0)[-2:]
    tied = (top2[1] - top2[0]) <= 1e-6 * np.maximum(top2[1], 1e-30)
    assert not (differ & ~tied).any(), f"{int((differ & ~tied).sum())} winners differ away from a weight tie"
    print(f"winners: {int(differ.sum())} differ, all at weight ties")

def test_golden_identity_remap(ctx, fg):
    """port3_0000.tif is the reference's own remap output of 0001.jpg under the identity pose."""
    d = np.load(os.path.join(GOLDEN, "s003_identity.npz"))
    out = ctx.remap(d["src"], fg.pose())
    assert np.array_equal(out, d["remap"])


def test_golden_aligned_output_file(ctx, fg, tmp_path):
    """fg_write_aligned_tiff (GPU remap + TIFF writer) reproduces the reference's aligned output
    port3_0000.tif byte for byte, under the name writeAlignedImageTask gives it."""
    import hashlib
    sig = json.load(open(os.path.join(GOLDEN, "s003_port3_0000_tiff.json")))
    d = np.load(os.path.join(GOLDEN, "s003_identity.npz"))
    path = ctx.write_aligned_tiff(d["src"], fg.pose(), None, str(tmp_path / "port3"), 0)
    assert os.path.basename(path) == "port3_0000.tif"
    raw = open(path, "rb").read()
    assert len(raw) == sig["size"] and hashlib.sha256(raw).hexdigest() == sig["sha256"]


@pytest.mark.parametrize("dtype,channels", [(np.uint8, 3), (np.uint16, 1)])
def test_aligned_output_is_the_remap(ctx, fg, tmp_path, dtype, channels):
    """Under a solved pose with an ROI: the written file is the TIFF of fg_remap's output."""
    meta = json.load(open(os.path.join(GOLDEN, "s003_pair_pose.json")))
    rng = np.random.default_rng(7 + channels)
    h, w = 97, 151
    shape = (h, w, channels) if channels > 1 else (h, w)
    src = rng.integers(0, np.iinfo(dtype).max + 1, shape).astype(dtype)
    p = fg.pose(meta["poses"][1])
    roi = (3, -2, w - 4, h + 1)
    path = ctx.write_aligned_tiff(src, p, roi, str(tmp_path / "al"), 12345)
    assert os.path.basename(path) == "al_12345.tif"
    ref = str(tmp_path / "direct.tif")
    fg.write_tiff(ref, ctx.remap(src, p, roi), extra=1)
    assert open(path, "rb").read() == open(ref, "rb").read()


POSES = {
    "identity": {},
    "similarity": dict(yaw=0.004, pitch=-0.003, roll=0.002, hfov_delta=0.2),
    "lens": dict(yaw=0.004, pitch=-0.003, roll=0.002, hfov_delta=0.2, radial_a=2e-3, radial_b=-1e-3,
                 radial_c=5e-4, center_shift_x=0.5, center_shift_y=-0.25),
    "shift_only": dict(center_shift_x=0.3, center_shift_y=0.1),
    "translation": dict(yaw=0.004, pitch=-0.003, roll=0.012, hfov_delta=0.15, trans_x=-0.05, trans_y=0.04,
                        trans_z=-0.015, translation_plane_yaw=0.02, translation_plane_pitch=-0.015,
                        radial_a=0.002, radial_b=-0.001, radial_c=0.0005, center_shift_x=1.0,
                        center_shift_y=-0.5),
}


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("channels", [3, 1])
@pytest.mark.parametrize("kind", list(POSES))
def test_remap_matches_oracle(ctx, fg, oracle, dtype, channels, kind):
    rng = np.random.default_rng(zlib.crc32(f"{kind}/{channels}".encode()))
    h, w = 131, 197
    maxv = 255 if dtype == np.uint8 else 65535
    shape = (h, w, channels) if channels > 1 else (h, w)
    src = rng.integers(0, maxv + 1, shape).astype(dtype)
    from oracle import make_pose
    roi = (-6, -4, w + 5, h + 3)  # partly outside the frame: exercises the support alpha border
    got = ctx.remap(src, fg.pose(POSES[kind]), roi)
    ref = oracle.remap(src, make_pose(**POSES[kind]), roi)
    assert got.shape == ref.shape
    _remap_close(got, ref)


def test_remap_real_solved_pose(ctx, fg, oracle):
    meta = json.load(open(os.path.join(GOLDEN, "s003_pair_pose.json")))
    d = np.load(os.path.join(GOLDEN, "s003_identity.npz"))
    from oracle import make_pose
    p = meta["poses"][1]
    got = ctx.remap(d["src"], fg.pose(p))
    ref = oracle.remap(d["src"], make_pose(**p))
    _remap_close(got, ref)


LENS_ROW_WALK = {
    # medium-size u16 lens stacks of the seeded generator (stack, frame): one sample of each rounded
    # the other way while the kernels formed the homography rows as products instead of walking them
    "1425x1259": (dict(frames=8, width=1425, height=1259, dtype="u16", lens=True, shift=3.474306049699964,
                       roll_deg=0.22336276936148092, scale=0.0007416607568523637, sigma_max=3.0), 1340, 7),
    "1618x892": (dict(frames=8, width=1618, height=892, dtype="u16", lens=True, shift=0.11441781908806714,
                      roll_deg=0.037218691316370164, scale=0.005976056069915496, sigma_max=3.0), 1341, 2),
}


@pytest.mark.parametrize("case", list(LENS_ROW_WALK))
def test_lens_row_walk_remap_exact(ctx, fg, oracle, case):
    """The lens path walks each output row adding the homography's first column once per pixel
    (src/remap.zig:397-437), so pixel x sees x roundings; the remap reproduces that sum exactly
    (fg_seqsum.h), and every sample equals the oracle's, which walks the row the same way."""
    import synth
    from oracle import make_pose
    kw, stack_id, k = LENS_ROW_WALK[case]
    spec = synth.config_spec(0, stack_id=stack_id, **kw)
    st = synth.Stack(spec)
    f = st.frame_numpy(k)
    opo = [make_pose(**p) for p in st.poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    got = ctx.remap(f, fg.pose(st.poses[k]), roi)
    ref = oracle.remap(f, opo[k], roi)
    assert np.array_equal(got, ref), f"{int((got != ref).sum())} samples differ"


# --------------------------------------------------------------------------
# whole stacks
# --------------------------------------------------------------------------

def _synthetic(spec_kw, lens=False):
    import synth
    spec = synth.config_spec(0, lens=lens, **spec_kw)
    st = synth.Stack(spec)
    frames = [st.frame_numpy(k) for k in range(spec.frames)]
    return spec, frames, st.poses


def _run_stack(fg, ctx, frames, poses, roi, **kw):
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    st = fg.SAMPLE_U8 if frames[0].dtype == np.uint8 else fg.SAMPLE_U16
    s = fg.Stack(ctx, len(frames), w, h, st, **kw)
    for f, p in zip(frames, poses):
        s.add_frame(f, fg.pose(p), roi)
    if s.needs_pass2_frames():
        for f, p in zip(frames, poses):
            s.pass2_frame(f, fg.pose(p), roi)
    out = s.finish()
    return s, out


CASES = {
    "small_u8": dict(frames=4, width=160, height=120),
    "odd_u8": dict(frames=3, width=99, height=67),
    "clamp_regime_u8": dict(frames=3, width=720, height=560),   # level 0 and 1 expand in the >512 regime
    "u16": dict(frames=3, width=131, height=97, dtype="u16"),
    "clamp_regime_u16": dict(frames=3, width=720, height=560, dtype="u16"),
    "host_ring_u8": dict(frames=11, width=160, height=120),       # more host frames than upload slots
    "tall": dict(frames=3, width=40, height=150),                # source levels one pixel wide
}


@pytest.mark.parametrize("case", list(CASES))
def test_stack_matches_oracle(ctx, fg, oracle, case):
    spec, frames, poses = _synthetic(CASES[case])
    from oracle import make_pose
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    s, out = _run_stack(fg, ctx, frames, poses, roi, debug_frame=1)
    ref, dbg = oracle.stack(frames, opo, roi, debug=True, frame_index=1)
    d = _img_close(out, ref)
    print(f"stack {case} roi={roi}: max {d.max()} mean {d.mean():.2e} differing {float((d != 0).mean()):.2e}")
    # intermediate buffers
    U = s.readback("union_support")
    assert np.array_equal(U, dbg["union_support"])
    S = s.readback("norm_sum")
    _level_close(S, dbg["norm_sum"], 1e-5)
    for k in range(len(frames)):
        _level_close(s.readback("weights", k), dbg["weights"][k], 1e-5)
    for l in range(len(s.levels())):
        _level_close(s.readback("acc_level", l), dbg["acc_levels"][l], 1e-5)
        _level_close(s.readback("mask_level", l), dbg["mask_levels"][l], 1e-5)
        _level_close(s.readback("gauss_level", l), dbg["gauss_levels"][l], 1e-5)
    s.close()


@pytest.mark.parametrize("dtype,method", [("u8", "pyramid"), ("u16", "pyramid"), ("u8", "hybrid"),
                                          ("u16", "softmask")])
def test_stack_lens_poses(ctx, fg, oracle, dtype, method):
    """Real solved stacks carry radial terms and centre shifts: the lens remap path (Newton
    undistortion, src/remap.zig:755-786) under every pass-1 variant."""
    spec, frames, poses = _synthetic(dict(frames=4, width=150, height=110, dtype=dtype), lens=True)
    from oracle import make_pose
    import oracle as om
    fm = {"pyramid": fg.METHOD_PYRAMID, "hybrid": fg.METHOD_HYBRID, "softmask": fg.METHOD_SOFTMASK}[method]
    o_m = {"pyramid": om.METHOD_PYRAMID, "hybrid": om.METHOD_HYBRID, "softmask": om.METHOD_SOFTMASK}[method]
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    s, out = _run_stack(fg, ctx, frames, poses, roi, method=fm)
    _img_close(out, oracle.stack(frames, opo, roi, method=o_m))
    s.close()


@pytest.mark.parametrize("sharpness", [0.35, 0.0, 1.0])
def test_hybrid_matches_oracle(ctx, fg, oracle, sharpness):
    spec, frames, poses = _synthetic(dict(frames=4, width=170, height=115))
    from oracle import make_pose
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    s, out = _run_stack(fg, ctx, frames, poses, roi, method=fg.METHOD_HYBRID, sharpness=sharpness)
    ref, dbg = oracle.stack(frames, opo, roi, method=1, sharpness=sharpness, debug=True)
    _img_close(s.readback("soft"), dbg["soft"])
    assert np.array_equal(s.readback("hard"), dbg["hard"])
    _img_close(out, ref)
    s.close()


@pytest.mark.parametrize("window,dtype", [(3, "u8"), (7, "u8"), (9, "u8"), (11, "u8"), (21, "u8"), (31, "u8"),
                                          (11, "u16"), (31, "u16")])
def test_contrast_window_sizes(ctx, fg, oracle, window, dtype):
    """--contrast-window-size 3..31 (src/fuse/config.zig:295-299): windows up to 7 use the fixed
    cross-lane sums, wider ones the generic sum, and from 11 on pass 1 runs 64-row tiles."""
    spec, frames, poses = _synthetic(dict(frames=3, width=120, height=90, dtype=dtype))
    from oracle import make_pose
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    for method, om in ((fg.METHOD_PYRAMID, 0), (fg.METHOD_HYBRID, 1)):
        s, out = _run_stack(fg, ctx, frames, poses, roi, window=window, method=method)
        _img_close(out, oracle.stack(frames, opo, roi, window=window, method=om))
        s.close()


@pytest.mark.parametrize("alpha", [False, True])
def test_aligned_frames(ctx, fg, oracle, alpha):
    """focus_fuse_zig path: pre-aligned frames, with or without an alpha channel."""
    rng = np.random.default_rng(5)
    spec, frames, poses = _synthetic(dict(frames=3, width=90, height=70))
    if alpha:
        frames = [np.concatenate([f, np.full(f.shape[:2] + (1,), 255, np.uint8)], -1) for f in frames]
        for f in frames:
            f[:3, :, 3] = 0
            f[:, -4:, 3] = rng.integers(0, 256, (f.shape[0], 4))
    extra = 1 if alpha else 0
    s = fg.Stack(ctx, len(frames), 90, 70)
    for f in frames:
        s.add_frame(f, None, None, extra=extra)
    out = s.finish()
    s.close()
    _img_close(out, oracle.fuse_frames(frames, extra=extra))


TINY = [(1, 1), (2, 1), (1, 2), (2, 2), (3, 3), (4, 3), (5, 5), (6, 2), (1, 9), (9, 1), (17, 3)]
TINY_METHODS = ["pyramid", "hybrid", "hardmask", "softmask"]


@pytest.mark.parametrize("method", TINY_METHODS)
@pytest.mark.parametrize("size", TINY, ids=lambda s: f"{s[0]}x{s[1]}")
def test_tiny_images(ctx, fg, oracle, size, method):
    """Images smaller than the contrast window (every weight 0, src/fuse/contrast.zig:113-115),
    one-pixel rows and columns, and a 1x1 image whose level 0 is its top level."""
    import oracle as om
    w, h = size
    rng = np.random.default_rng(w * 100 + h)
    frames = [rng.integers(0, 256, (h, w, 3)).astype(np.uint8) for _ in range(3)]
    fm = {"pyramid": fg.METHOD_PYRAMID, "hybrid": fg.METHOD_HYBRID, "hardmask": fg.METHOD_HARDMASK,
          "softmask": fg.METHOD_SOFTMASK}[method]
    o_m = {"pyramid": om.METHOD_PYRAMID, "hybrid": om.METHOD_HYBRID, "hardmask": om.METHOD_HARDMASK,
           "softmask": om.METHOD_SOFTMASK}[method]
    s = fg.Stack(ctx, len(frames), w, h, method=fm)
    for f in frames:
        s.add_frame(f, None, None)
    out = s.finish()
    s.close()
    ref = oracle.fuse_frames(frames, extra=0, method=o_m)
    if method == "hardmask":
        assert np.array_equal(out, ref)
    else:
        _img_close(out, ref)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
@pytest.mark.parametrize("size", TINY, ids=lambda s: f"{s[0]}x{s[1]}")
def test_tiny_sources_remapped(ctx, fg, oracle, size, dtype):
    """The remap path on tiny sources: the identity pose over the full frame, whose border
    samples have partial support."""
    from oracle import make_pose
    w, h = size
    rng = np.random.default_rng(w * 100 + h + 7)
    hi = 256 if dtype == np.uint8 else 65536
    frames = [rng.integers(0, hi, (h, w, 3)).astype(dtype) for _ in range(3)]
    poses = [dict(), dict(roll=0.01), dict(yaw=-0.004, hfov_delta=0.3)]
    roi = (0, 0, w, h)
    s, out = _run_stack(fg, ctx, frames, poses, roi)
    s.close()
    _img_close(out, oracle.stack(frames, [make_pose(**p) for p in poses], roi))


@pytest.mark.parametrize("levels", [0, 6])
def test_cpu_runnable_config_matches_oracle(ctx, fg, oracle, levels):
    """The reference's CPU-runnable case at its own shape: 8 seeded 1024x768 u8 frames under
    similarity poses, remap + pyramid-contrast, at full depth (11 levels) and capped at the 6
    levels its configuration names; the fused image and every intermediate level."""
    spec, frames, poses = _synthetic(dict())
    assert (spec.frames, spec.width, spec.height) == (8, 1024, 768)
    from oracle import make_pose
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    s, out = _run_stack(fg, ctx, frames, poses, roi, max_levels=levels, debug_frame=3)
    ref, dbg = oracle.stack(frames, opo, roi, max_levels=levels, debug=True, frame_index=3)
    assert len(s.levels()) == (11 if levels == 0 else levels)
    d = _img_close(out, ref)
    print(f"8 x 1024x768 u8 levels={levels} roi={roi}: max {d.max()} mean {d.mean():.2e}")
    _level_close(s.readback("norm_sum"), dbg["norm_sum"], 1e-5)
    for l in range(len(s.levels())):
        _level_close(s.readback("acc_level", l), dbg["acc_levels"][l], 1e-5)
        _level_close(s.readback("mask_level", l), dbg["mask_levels"][l], 1e-5)
        _level_close(s.readback("gauss_level", l), dbg["gauss_levels"][l], 1e-5)
    s.close()


def test_levels_cap(ctx, fg, oracle):
    spec, frames, poses = _synthetic(dict(frames=3, width=128, height=96))
    from oracle import make_pose
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    s, out = _run_stack(fg, ctx, frames, poses, roi, max_levels=4)
    assert len(s.levels()) == 4
    _img_close(out, oracle.stack(frames, opo, roi, max_levels=4))
    s.close()


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_translation_poses_stack(ctx, fg, oracle, dtype):
    """Poses with translation terms (src/optimize.zig:4009-4043): every pixel of pass 1 takes the
    exact remap path instead of the homography fast path."""
    from oracle import make_pose
    kw = dict(frames=3, width=140, height=100) if dtype == np.uint8 else dict(frames=3, width=140, height=100,
                                                                               dtype="u16")
    spec, frames, poses = _synthetic(kw)
    tr = [dict(), dict(trans_x=0.02, trans_y=-0.015, trans_z=0.01, translation_plane_yaw=0.01),
          dict(trans_x=-0.01, trans_y=0.02, trans_z=-0.012, translation_plane_pitch=-0.01)]
    poses = [dict(p, **t) for p, t in zip(poses, tr)]
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    s, out = _run_stack(fg, ctx, frames, poses, roi)
    s.close()
    _img_close(out, oracle.stack(frames, opo, roi))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_unaligned_device_frames(ctx, fg, dtype):
    """Device frames that start off a dword boundary and end at the last byte of their
    allocation: pass 1 stages the source through a dword-aligned view, assembling the dwords
    that straddle either end of the frame from its own bytes.  The result equals the stack
    fed from aligned buffers."""
    import ctypes as C
    kw = dict(frames=3, width=150, height=110) if dtype == np.uint8 else dict(frames=3, width=131, height=97,
                                                                               dtype="u16")
    spec, frames, poses = _synthetic(kw)
    roi = fg.common_overlap_roi([(spec.width, spec.height)] * 3, [fg.pose(p) for p in poses])
    _, ref = _run_stack(fg, ctx, frames, poses, roi)
    st = fg.SAMPLE_U8 if dtype == np.uint8 else fg.SAMPLE_U16
    sb = np.dtype(dtype).itemsize
    desc = fg.ImageDesc(spec.width, spec.height, 3, 0, st)
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    s = fg.Stack(ctx, 3, w, h, st)
    allocs = []
    for k, (f, p) in enumerate(zip(frames, poses)):
        off = (k + 1) * sb                      # 1, 2, 3 bytes (u8) or 2, 4, 6 (u16, sample-aligned)
        f = np.ascontiguousarray(f)
        d = ctx.alloc(f.nbytes + off)           # the frame's last byte is the allocation's last byte
        allocs.append(d)
        fg.check(fg.lib().fg_copy_to_device(ctx.h, C.c_void_p(d + off), fg._p(f), f.nbytes))
        s.add_frame_device(d + off, desc, fg.pose(p), roi)
    out = s.finish()
    s.close()
    for d in allocs:
        ctx.free(d)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("method,dtype", [("pyramid", "u8"), ("hybrid", "u8"), ("hybrid", "u16")])
def test_graph_replays_equal_eager(ctx, fg, method, dtype):
    """A cached stack finished again runs pass 2 + collapse as a HIP graph captured on its second
    finish; with profiling on it launches the same kernels eagerly.  Two frame sets alternate through one stack
    (eager pass 1, reset and synchronisation between the runs), so a replay that kept stale data
    or a captured clear that misbehaved on later replays would show.  The u16 hybrid takes its
    winners from the cached weights after pass 1."""
    from oracle import make_pose
    m = fg.METHOD_PYRAMID if method == "pyramid" else fg.METHOD_HYBRID
    spec, fa, poses = _synthetic(dict(frames=4, width=160, height=120, dtype=dtype))
    fb = [np.ascontiguousarray(f[:, ::-1]) for f in fa[::-1]]   # a different stack of the same geometry
    roi = fg.common_overlap_roi([(spec.width, spec.height)] * 4, [fg.pose(p) for p in poses])
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    s = fg.Stack(ctx, 4, w, h, fg.SAMPLE_U8 if dtype == "u8" else fg.SAMPLE_U16, method=m)

    def run(frames):
        s.reset()
        for f, p in zip(frames, poses):
            s.add_frame(f, fg.pose(p), roi)
        out = s.finish()
        s.sync()
        return out

    s.set_profiling(True)                       # eager launches
    eager = {"a": run(fa), "b": run(fb)}
    s.set_profiling(False)                      # captured on the next finish, replayed after
    assert not np.array_equal(eager["a"], eager["b"])
    for i, key in enumerate("ababab"):
        out = run(fa if key == "a" else fb)
        assert np.array_equal(out, eager[key]), f"run {i} ({key}) differs from the eager result"
    s.close()


@pytest.mark.parametrize("dtype", ["u8", "u16"])
def test_hybrid_host_frames_several_launches(ctx, fg, oracle, dtype):
    """Hybrid fed from host buffers: pass 1 runs once per 8 uploaded frames, and each launch's
    winners are resolved against the hard image of the earlier launches (u8: kept as frame slots of
    the cache inside pass 1; u16: taken from the cached weights by the winners kernel after it)."""
    from oracle import make_pose, METHOD_HYBRID
    spec, frames, poses = _synthetic(dict(frames=11, width=160, height=120, dtype=dtype))
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    s, out = _run_stack(fg, ctx, frames, poses, roi, method=fg.METHOD_HYBRID, debug_frame=1)
    ref, dbg = oracle.stack(frames, opo, roi, method=METHOD_HYBRID, debug=True, frame_index=1)
    _img_close(out, ref)
    assert np.array_equal(s.readback("hard"), dbg["hard"])
    s.close()


@pytest.mark.parametrize("levels", [3, 4, 6, 7])
def test_hybrid_levels_cap(ctx, fg, oracle, levels):
    """Hybrid with the pyramid capped: at 6 levels or fewer the hard range [2, 6) reaches the top
    level (the gaussian), which mergeHybridImages mixes like any other (src/fuse/pyramid.zig:288-298)."""
    spec, frames, poses = _synthetic(dict(frames=3, width=128, height=96))
    from oracle import make_pose, METHOD_HYBRID
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    s, out = _run_stack(fg, ctx, frames, poses, roi, max_levels=levels, method=fg.METHOD_HYBRID)
    assert len(s.levels()) == levels
    _img_close(out, oracle.stack(frames, opo, roi, max_levels=levels, method=METHOD_HYBRID))
    s.close()


@pytest.mark.parametrize("method", ["pyramid", "hybrid"])
def test_uncached_second_pass_equals_cached(ctx, fg, oracle, method):
    """With a memory budget too small for the frame cache the caller feeds the frames again.  For
    the hybrid, the cached stack keeps each winner as a frame slot and reads its colour back from
    the frame cache; the uncached one keeps the colour itself."""
    m = fg.METHOD_PYRAMID if method == "pyramid" else fg.METHOD_HYBRID
    spec, frames, poses = _synthetic(dict(frames=3, width=100, height=80))
    from oracle import make_pose
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    s1, out1 = _run_stack(fg, ctx, frames, poses, roi, method=m)
    s2, out2 = _run_stack(fg, ctx, frames, poses, roi, memory_fraction=1e-9, method=m)
    assert not s1.needs_pass2_frames() and s2.needs_pass2_frames()
    assert np.array_equal(out1, out2)
    if method == "hybrid":
        assert np.array_equal(s1.readback("hard"), s2.readback("hard"))
    # reset and reused with another frame set: pass 2 starts its sums at frame 0 without reading
    # the accumulator (reset leaves it as the last finish did), so a stale sum would show here
    flipped = [np.ascontiguousarray(f[::-1, ::-1]) for f in frames[::-1]]
    s3, fresh = _run_stack(fg, ctx, flipped, poses, roi, method=m)
    s3.close()
    assert not np.array_equal(fresh, out1)
    for st in (s1, s2):
        st.reset()
        assert not st.readback("acc_level", 0).any()     # nothing accumulated yet reads as zeros
        for f, p in zip(flipped, poses):
            st.add_frame(f, fg.pose(p), roi)
        if st.needs_pass2_frames():
            for f, p in zip(flipped, poses):
                st.pass2_frame(f, fg.pose(p), roi)
        assert np.array_equal(st.finish(), fresh)
    s1.close()
    s2.close()


def test_single_frame_and_identical_frames(ctx, fg):
    spec, frames, poses = _synthetic(dict(frames=2, width=64, height=48))
    f = frames[0]
    for n in (1, 3):
        s = fg.Stack(ctx, n, 64, 48)
        for _ in range(n):
            s.add_frame(f)
        out = s.finish()
        s.close()
        assert np.array_equal(out, f)


def test_errors_surface_as_status_codes(ctx, fg):
    s = fg.Stack(ctx, 2, 32, 24)
    with pytest.raises(fg.FgError) as e:
        s.add_frame(np.zeros((20, 32, 3), np.uint8))
    assert e.value.code == -4  # FG_ERR_SIZE_MISMATCH
    with pytest.raises(fg.FgError) as e:
        s.add_frame(np.zeros((24, 32, 3), np.uint16))
    assert e.value.code == -5  # FG_ERR_FORMAT_MISMATCH
    with pytest.raises(fg.FgError) as e:
        s.finish()
    assert e.value.code == -8  # FG_ERR_STATE
    # the kernels address frames and levels with 32-bit byte offsets: larger ones are refused
    # before anything runs
    import torch
    dummy = torch.zeros(16, dtype=torch.uint8, device="cuda:0")
    with pytest.raises(fg.FgError) as e:
        s.add_frame_device(dummy.data_ptr(), fg.ImageDesc(40000, 40000, 3, 0, fg.SAMPLE_U8),
                           fg.pose(), (0, 0, 32, 24))
    assert e.value.code == -6  # FG_ERR_UNSUPPORTED
    s.close()
    with pytest.raises(fg.FgError):
        fg.Stack(ctx, 2, 32, 24, window=4)
    with pytest.raises(fg.FgError) as e:
        fg.Stack(ctx, 2, 16384, 16384)
    assert e.value.code == -6  # FG_ERR_UNSUPPORTED: 2^28 pixels


SINGLE_CASES = {
    "hardmask_u8": ("hardmask", dict(frames=5, width=150, height=110)),
    "hardmask_groups_u8": ("hardmask", dict(frames=11, width=97, height=70)),   # winners carried across launches
    "softmask_u8": ("softmask", dict(frames=5, width=150, height=110)),
    "softmask_groups_u8": ("softmask", dict(frames=11, width=97, height=70)),   # more frames than one group
    "softmask_u16": ("softmask", dict(frames=3, width=120, height=90, dtype="u16")),
    "hardmask_u16": ("hardmask", dict(frames=3, width=120, height=90, dtype="u16")),
}


def _sweep_case(seed):
    """A seeded random stack: size, frame count, sample type, method, contrast window, pose
    magnitudes and the lens path drawn together, so the sweep crosses combinations that the
    hand-picked cases above do not."""
    rng = np.random.default_rng(90210 + seed)
    small = rng.random() < 0.3
    w = int(rng.integers(2, 40)) if small else int(rng.integers(40, 421))   # the generator needs 2+
    h = int(rng.integers(2, 40)) if small else int(rng.integers(40, 421))
    return dict(seed=seed, w=w, h=h, frames=int(rng.integers(1, 7)),
                dtype=("u8", "u16")[(seed // 4) % 2],
                method=("pyramid", "hybrid", "hardmask", "softmask")[seed % 4],
                window=int(rng.choice([3, 5, 5, 7, 9, 11])), lens=bool(rng.random() < 0.3),
                shift=float(rng.uniform(0.0, min(3.0, max(w, h) / 10.0))),
                roll=float(rng.uniform(0.0, 0.3)), scale=float(rng.uniform(0.0, 0.01)))


@pytest.mark.parametrize("seed", range(256))
def test_random_stacks_match_oracle(ctx, fg, oracle, seed):
    """Seeded sweep over whole stacks through the C ABI, each against the oracle with the same
    options; the case is printed first so a failure names it."""
    import synth
    import oracle as om
    c = _sweep_case(seed)
    print(f"sweep case {c}", flush=True)
    spec = synth.config_spec(0, stack_id=1000 + seed, frames=c["frames"], width=c["w"], height=c["h"],
                             dtype=c["dtype"], lens=c["lens"], shift=c["shift"], roll_deg=c["roll"],
                             scale=c["scale"], sigma_max=3.0)
    st = synth.Stack(spec)
    frames = [st.frame_numpy(k) for k in range(spec.frames)]
    poses = st.poses
    from oracle import make_pose
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    if roi is None or roi[2] - roi[0] < 1 or roi[3] - roi[1] < 1:  # tiny frames shifted apart
        roi = (0, 0, spec.width, spec.height)
    fm = {"pyramid": fg.METHOD_PYRAMID, "hybrid": fg.METHOD_HYBRID, "hardmask": fg.METHOD_HARDMASK,
          "softmask": fg.METHOD_SOFTMASK}[c["method"]]
    o_m = {"pyramid": om.METHOD_PYRAMID, "hybrid": om.METHOD_HYBRID, "hardmask": om.METHOD_HARDMASK,
           "softmask": om.METHOD_SOFTMASK}[c["method"]]
    s, out = _run_stack(fg, ctx, frames, poses, roi, method=fm, window=c["window"])
    s.close()
    ref = oracle.stack(frames, opo, roi, method=o_m, window=c["window"])
    assert out.shape == ref.shape
    if c["method"] == "hardmask":
        weights = None
        if c["dtype"] == "u16":
            _, dbg = oracle.stack(frames, opo, roi, method=om.METHOD_PYRAMID, window=c["window"], debug=True)
            weights = dbg["weights"]
        _winners_exact_or_tied(out, ref, weights, 2 if c["dtype"] == "u16" else 1)
    else:
        _img_close(out, ref)


@pytest.mark.parametrize("case", list(SINGLE_CASES))
def test_single_pass_methods_match_oracle(ctx, fg, oracle, case):
    """hardmask-contrast and softmask-contrast (src/stack/pipeline.zig:245-312)."""
    from oracle import make_pose, METHOD_HARDMASK, METHOD_SOFTMASK
    method, kw = SINGLE_CASES[case]
    spec, frames, poses = _synthetic(kw)
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    fm, om = (fg.METHOD_HARDMASK, METHOD_HARDMASK) if method == "hardmask" else (fg.METHOD_SOFTMASK, METHOD_SOFTMASK)
    s, out = _run_stack(fg, ctx, frames, poses, roi, method=fm)
    assert not s.needs_pass2_frames()
    ref = oracle.stack(frames, opo, roi, method=om)
    if method == "hardmask":
        weights = None
        if frames[0].dtype == np.uint16:
            from oracle import METHOD_PYRAMID
            weights = oracle.stack(frames, opo, roi, method=METHOD_PYRAMID, debug=True)[1]["weights"]
        _winners_exact_or_tied(out, ref, weights, frames[0].dtype.itemsize)
    else:
        _img_close(out, ref)
    s.close()


def test_softmask_aligned_frames(ctx, fg, oracle):
    from oracle import METHOD_SOFTMASK
    g = np.random.default_rng(5)
    frames = [g.integers(0, 256, (60, 80, 4)).astype(np.uint8) for _ in range(4)]
    for f in frames:
        f[..., 3] = 255
        f[:3, :, 3] = 0
    s = fg.Stack(ctx, 4, 80, 60, fg.SAMPLE_U8, method=fg.METHOD_SOFTMASK)
    for f in frames:
        s.add_frame(f, extra=1)
    out = s.finish()
    _img_close(out, oracle.fuse_frames(frames, method=METHOD_SOFTMASK))
    s.close()


def test_pinned_host_frames_match_pageable(ctx, fg, oracle):
    """Frames in fg_host_alloc memory go through the same upload path and give the same image."""
    spec, frames, poses = _synthetic(dict(frames=3, width=140, height=100))
    from oracle import make_pose
    opo = [make_pose(**p) for p in poses]
    roi = oracle.common_overlap_roi([(spec.width, spec.height)] * spec.frames, opo)
    _, out1 = _run_stack(fg, ctx, frames, poses, roi)
    pinned = []
    for f in frames:
        pa = fg.PinnedArray(f.shape, f.dtype)
        pa.array[...] = f
        pinned.append(pa)
    _, out2 = _run_stack(fg, ctx, [p.array for p in pinned], poses, roi)
    for p in pinned:
        p.close()
    assert np.array_equal(out1, out2)


def test_host_frames_in_many_chunks_match_device_frames(ctx, fg):
    """Host-fed frames larger than one 16 MiB upload chunk (3000x2000 u16: 36 MB, three chunks, the
    last one partial), more frames than upload slots, from pageable memory (threaded copies through
    the library's pinned chunks) and from page-locked memory (copied directly): the same image as
    the same frames fed from device memory."""
    import torch
    import synth
    spec = synth.config_spec(0, stack_id=77, frames=10, width=3000, height=2000, shift=3.0 * 3000 / 1024.0,
                             dtype="u16")
    st = synth.Stack(spec, device=torch.device("cuda:0"))
    dev, host = [], []
    for k in range(spec.frames):
        f = st.frame(k)
        host.append(f.cpu().numpy().astype(np.uint16))
        dev.append((f - (f > 32767).to(torch.int32) * 65536).to(torch.int16).contiguous())
    poses = [fg.pose(p) for p in st.poses]
    roi = fg.common_overlap_roi([(spec.width, spec.height)] * spec.frames, poses)
    w, h = roi[2] - roi[0], roi[3] - roi[1]
    torch.cuda.synchronize()
    s = fg.Stack(ctx, spec.frames, w, h, fg.SAMPLE_U16)
    d = fg.ImageDesc(spec.width, spec.height, 3, 0, fg.SAMPLE_U16)
    for f, p in zip(dev, poses):
        s.add_frame_device(f.data_ptr(), d, p, roi)
    ref = s.finish()
    s.close()
    print("3000x2000 u16 x 10, roi", roi, flush=True)
    _, out = _run_stack(fg, ctx, host, st.poses, roi)
    assert np.array_equal(out, ref), f"pageable: {int((out != ref).sum())} samples differ"
    pinned = []
    for f in host:
        pa = fg.PinnedArray(f.shape, f.dtype)
        pa.array[...] = f
        pinned.append(pa)
    _, out2 = _run_stack(fg, ctx, [p.array for p in pinned], st.poses, roi)
    for p in pinned:
        p.close()
    assert np.array_equal(out2, ref), f"page-locked: {int((out2 != ref).sum())} samples differ"
