"""Pins the CPU oracle before it is trusted to check the HIP path.

Three kinds of checks, each chosen so that a plausible slip in the oracle
(dropped term, wrong index, transposed operand, wrong divisor) fails one:
  * the reference's own known-answer tests and golden data files;
  * closed forms / invariants the mathematics fixes;
  * brute-force restatements written differently (vectorised numpy) on tiny inputs.
Parts no check pins at the ulp level (libm-level rounding of the pose trigonometry, the fused
end-to-end image, which the reference holds no fixture for) stay "parity unpinned".
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import Oracle, make_pose, level_dims

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --------------------------------------------------------------------------
# reference known answers
# --------------------------------------------------------------------------

def test_identity_remap_gray_ramp(oracle):
    # src/remap.zig:1018-1057
    src = np.arange(1, 26, dtype=np.uint8).reshape(5, 5)
    out = oracle.remap(src, make_pose())
    assert out.shape == (5, 5, 2)
    assert out[2, 2, 0] == 13 and out[2, 2, 1] == 255
    assert out[3, 2, 0] == 18 and out[3, 2, 1] == 255
    assert out[2, 3, 0] == 14 and out[2, 3, 1] == 255


def test_overlap_roi_known_answers(oracle):
    # src/remap.zig:1059-1078 and :1080-1093
    roi = oracle.common_overlap_roi([(10, 8), (10, 8)], [make_pose(), make_pose()])
    assert roi is not None
    l, t, r, b = roi
    assert l >= 0 and t >= 0 and r <= 10 and b <= 8 and r > l and b > t
    roi = oracle.common_overlap_roi([(10, 8)], [make_pose(trans_z=-0.1)])
    assert roi is not None and roi[2] > roi[0] and roi[3] > roi[1]


def test_gray_average(oracle):
    # src/fuse/grayscale.zig:59-77
    img = np.array([[[0, 255, 255]]], np.uint8)
    g = oracle.gray(img, extra=0)
    assert abs(float(g[0, 0]) - 2.0 / 3.0) < 1e-6


def test_sample_std(oracle):
    # src/fuse/contrast.zig:297-306
    v = [0.0, 1.0, 1.0, 2.0]
    assert abs(oracle.sample_std(sum(v), sum(x * x for x in v), 4) - 0.8164966) < 1e-6
    assert oracle.sample_std(3.0, 9.0, 1) == 0.0


def test_normalize_fallback(oracle):
    # src/fuse/pyramid.zig:1858-1863
    out = oracle.normalize(np.array([1, 2], np.float32), np.array([2, 0], np.float32), 4)
    assert np.allclose(out, [0.5, 0.25], atol=1e-6)


def test_uniform_reduce_expand(oracle):
    # src/fuse/pyramid.zig:1865-1884
    src = np.tile(np.array([10, 20, 30], np.float32), (4, 4, 1))
    red = oracle.reduce_clamped(src)
    exp = oracle.expand(red, 4, 4)
    assert np.abs(exp - src).max() <= 1e-5


def test_winner_known_answers(oracle):
    # src/fuse/blend.zig:316-349
    img = np.array([[[1, 2, 3], [4, 5, 6]]], np.uint8)
    best = np.full((1, 2), -np.inf, np.float32)
    out = np.zeros((1, 2, 3), np.uint8)
    oracle.update_winners(img, np.array([[0.1, 0.2]], np.float32), best, out, extra=0)
    assert out.ravel().tolist() == [1, 2, 3, 4, 5, 6]
    # src/fuse/blend.zig:351-401: alpha discounts the weight
    best = np.full((1, 1), -np.inf, np.float32)
    out = np.zeros((1, 1, 3), np.uint8)
    oracle.update_winners(np.array([[[200, 10, 10, 32]]], np.uint8), np.array([[1.0]], np.float32), best, out)
    oracle.update_winners(np.array([[[10, 200, 10, 255]]], np.uint8), np.array([[0.2]], np.float32), best, out)
    assert out.ravel().tolist() == [10, 200, 10]


def test_transform_round_trip(oracle):
    # src/optimize.zig:4696-4719
    pose = make_pose(yaw=0.004, pitch=-0.003, roll=0.012, hfov_delta=0.15, trans_x=-3.2, trans_y=2.4,
                     trans_z=-0.015, translation_plane_yaw=0.02, translation_plane_pitch=-0.015,
                     radial_a=0.02, radial_b=-0.01, radial_c=0.005, center_shift_x=4.0, center_shift_y=-2.5)
    mx, my = oracle.transform_point(pose, 400, 300, 120.0, 80.0)
    rx, ry = oracle.inverse_point(pose, 400, 300, mx, my)
    assert abs(rx - 120.0) < 0.05 and abs(ry - 80.0) < 0.05


def test_level_counts(oracle):
    # computeLevelCount src/fuse/pyramid.zig:163-176: down to 1x1
    assert oracle.level_count(1024, 768) == 11
    assert oracle.level_count(7952, 5304) == 14
    assert oracle.level_count(1, 1) == 1
    assert oracle.level_count(2, 1) == 2
    assert len(level_dims(97, 65)) == oracle.level_count(97, 65)


# --------------------------------------------------------------------------
# golden data files of the reference
# --------------------------------------------------------------------------

def test_golden_identity_remap_matches_reference_output(oracle):
    """port3_0000.tif is the reference's remap of 0001.jpg under frame 0's identity pose."""
    d = np.load(os.path.join(GOLDEN, "s003_identity.npz"))
    out = oracle.remap(d["src"], make_pose())
    assert np.array_equal(out, d["remap"])


def test_real_solved_pose_and_roll_direction(oracle):
    """A pose solved by the reference (port_pair_m.pto) is basic rectilinear, so remap takes the
    cached legacy chain.  That chain uses setMatrix(-pitch, 0, roll) (src/optimize.zig:3571-3572)
    while the forward transformPoint uses setMatrix(+pitch, 0, roll) (:3676): pitch, yaw and hfov
    terms cancel in inverse(forward(x)), roll is applied twice (rotation by 2*roll)."""
    meta = json.load(open(os.path.join(GOLDEN, "s003_pair_pose.json")))
    pose = make_pose(**meta["poses"][1])
    w, h = meta["width"], meta["height"]
    assert oracle.inverse_cache(pose, w, h).basic_rectilinear == 1
    for kw in (dict(pitch=0.01), dict(yaw=0.01), dict(hfov_delta=0.5)):
        p = make_pose(**kw)
        mx, my = oracle.transform_point(p, w, h, 700.0, 40.0)
        rx, ry = oracle.inverse_point(p, w, h, mx, my)
        assert abs(rx - 700.0) < 1e-9 and abs(ry - 40.0) < 1e-9
    roll = 0.002
    p = make_pose(roll=roll)
    cx, cy = (w - 1) / 2, (h - 1) / 2
    for (x, y) in [(700.0, 40.0), (10.0, 500.0)]:
        mx, my = oracle.transform_point(p, w, h, x, y)
        rx, ry = oracle.inverse_point(p, w, h, mx, my)
        ang = math.atan2(ry - cy, rx - cx) - math.atan2(y - cy, x - cx)
        assert abs(abs(ang) - 2 * roll) < 1e-4 * roll + 1e-9


# --------------------------------------------------------------------------
# closed forms and invariants
# --------------------------------------------------------------------------

def _homography_inverse(cache, x, y):
    m = np.array(cache.pitch_roll[:]).reshape(3, 3)
    cy, sy = math.cos(cache.yaw), math.sin(cache.yaw)
    ry = np.array([[cy, 0, -sy], [0, 1, 0], [sy, 0, cy]])
    t = (m.T @ ry) @ np.array([x - cache.dest_center_x, y - cache.dest_center_y, cache.pano_distance])
    return (cache.src_center_x + cache.image_focal * t[0] / t[2],
            cache.src_center_y + cache.image_focal * t[1] / t[2])


@pytest.mark.parametrize("w,h", [(7952, 5304), (1024, 768), (97, 65)])
def test_legacy_chain_is_a_homography(oracle, w, h):
    """The basic-rectilinear inverse chain (erect->rotate->sphereTP->persp->rect) is the pinhole
    rotation homography img_f * (R p).xy / (R p).z with R = pitch_roll^T * yawMatrix(yaw)."""
    rng = np.random.default_rng(1)
    for _ in range(4):
        pose = make_pose(yaw=rng.uniform(-0.05, 0.05), pitch=rng.uniform(-0.05, 0.05),
                         roll=rng.uniform(-0.1, 0.1), hfov_delta=rng.uniform(-2, 2))
        c = oracle.inverse_cache(pose, w, h)
        assert c.basic_rectilinear == 1
        for _ in range(300):
            x, y = rng.uniform(-20, w + 20), rng.uniform(-20, h + 20)
            sx, sy = oracle.inverse_point(pose, w, h, x, y)
            hx, hy = _homography_inverse(c, x, y)
            assert abs(sx - hx) < 1e-9 and abs(sy - hy) < 1e-9


def test_inverse_is_inverse_of_forward_lens_and_translation(oracle):
    """Off the basic path (lens or translation terms) forward and cached inverse agree."""
    rng = np.random.default_rng(2)
    w, h = 640, 480
    poses = [
        make_pose(yaw=0.01, pitch=-0.02, roll=0.03, hfov_delta=0.5, radial_a=1e-3, radial_b=-5e-4,
                  radial_c=2e-4, center_shift_x=0.4, center_shift_y=-0.3),
        make_pose(yaw=0.004, pitch=-0.003, roll=0.012, trans_x=-0.2, trans_y=0.1, trans_z=-0.01),
    ]
    for pose in poses:
        for _ in range(50):
            x, y = rng.uniform(0, w - 1), rng.uniform(0, h - 1)
            mx, my = oracle.transform_point(pose, w, h, x, y)
            rx, ry = oracle.inverse_point(pose, w, h, mx, my)
            assert abs(rx - x) < 1e-3 and abs(ry - y) < 1e-3


def _bilinear_brute(src, xs, ys, maxv):
    """Support-alpha bilinear restated: taps outside the image drop out and the
    remaining weights are renormalised; alpha = round(in-bounds weight * maxv)."""
    h, w, ch = src.shape
    out = np.zeros((len(xs), ch + 1), np.int64)
    for i, (x, y) in enumerate(zip(xs, ys)):
        x0, y0 = math.floor(x), math.floor(y)
        fx, fy = x - x0, y - y0
        taps = [(x0, y0, (1 - fx) * (1 - fy)), (x0 + 1, y0, fx * (1 - fy)), (x0, y0 + 1, (1 - fx) * fy),
                (x0 + 1, y0 + 1, fx * fy)]
        inside = [(tx, ty, wt) for (tx, ty, wt) in taps if 0 <= tx < w and 0 <= ty < h]
        s = sum(wt for _, _, wt in inside)
        if s <= 0:
            continue
        for c in range(ch):
            v = sum(float(src[ty, tx, c]) * wt for tx, ty, wt in inside)
            out[i, c] = int(math.floor(v / s + 0.5))
        out[i, ch] = maxv if len(inside) == 4 else int(math.floor(s * maxv + 0.5))
    return out


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_remap_bilinear_support_brute_force(oracle, dtype):
    rng = np.random.default_rng(3)
    maxv = 255 if dtype == np.uint8 else 65535
    src = rng.integers(0, maxv + 1, (9, 13, 3)).astype(dtype)
    pose = make_pose(yaw=0.03, pitch=-0.02, roll=0.05, hfov_delta=1.5)
    out = oracle.remap(src, pose, roi=(-3, -2, 16, 12))
    xs, ys = [], []
    for oy in range(-2, 12):
        for ox in range(-3, 16):
            sx, sy = oracle.inverse_point(pose, 13, 9, float(ox), float(oy))
            xs.append(sx)
            ys.append(sy)
    ref = _bilinear_brute(src, xs, ys, maxv).reshape(14, 19, 4)
    assert np.array_equal(out.astype(np.int64), ref)
    assert (out[..., 3] == 0).any() and (out[..., 3] == maxv).any()  # both borders exercised


def test_remap_lens_path_brute_force(oracle):
    """Fast lens path (incremental homography + Newton undistort) against the per-pixel cached inverse."""
    rng = np.random.default_rng(4)
    src = rng.integers(0, 256, (40, 60, 3)).astype(np.uint8)
    pose = make_pose(yaw=0.01, pitch=0.004, roll=-0.02, hfov_delta=0.3, radial_a=2e-3, radial_b=-1e-3,
                     radial_c=5e-4, center_shift_x=0.5, center_shift_y=-0.25)
    out = oracle.remap(src, pose)
    xs, ys = [], []
    for oy in range(40):
        for ox in range(60):
            sx, sy = oracle.inverse_point(pose, 60, 40, float(ox), float(oy))
            xs.append(sx)
            ys.append(sy)
    ref = _bilinear_brute(src, xs, ys, 255).reshape(40, 60, 4)
    assert np.abs(out.astype(int) - ref).max() <= 1
    assert (out.astype(int) != ref).mean() < 1e-3


def _contrast_brute(gray, sup, win):
    h, w = gray.shape
    b = win // 2
    out = np.zeros_like(gray, dtype=np.float64)
    for y in range(b, h - b):
        for x in range(b, w - b):
            if sup[y, x] <= 0:
                continue
            blk = gray[y - b:y + b + 1, x - b:x + b + 1].astype(np.float64)
            s = sup[y - b:y + b + 1, x - b:x + b + 1] > 0
            v = blk[s]
            out[y, x] = v.std(ddof=1) if v.size > 1 else 0.0
    return out


@pytest.mark.parametrize("win", [3, 5, 7])
def test_contrast_brute_force(oracle64, win):
    rng = np.random.default_rng(5)
    gray = rng.random((17, 23))
    sup = (rng.random((17, 23)) > 0.2).astype(np.float64)
    got = oracle64.contrast(gray, sup, win)
    ref = _contrast_brute(gray, sup, win)
    assert np.abs(got - ref).max() < 1e-9
    # image smaller than the window -> all zero (src/fuse/contrast.zig:113-115)
    assert not oracle64.contrast(gray[:win - 1], sup[:win - 1], win).any()


def _sep(v, k):
    return np.convolve(v, k, mode="same")


def test_reduce_support_is_normalised_convolution(oracle64):
    """Binary support + separable kernel: reduce == sepconv(s*v)/sepconv(s) at even sites."""
    rng = np.random.default_rng(6)
    for (h, w) in [(9, 12), (10, 7), (1, 5), (6, 1)]:
        v = rng.random((h, w, 3))
        s = (rng.random((h, w)) > 0.3).astype(np.float64)
        got, gs = oracle64.reduce_support(v, s)
        k = np.array([1, 4, 6, 4, 1], np.float64)
        pad = np.pad(s, 2)
        num = np.pad(v * s[..., None], ((2, 2), (2, 2), (0, 0)))
        dh, dw = gs.shape
        for dy in range(dh):
            for dx in range(dw):
                ws = (k[:, None] * k[None, :] * pad[2 * dy:2 * dy + 5, 2 * dx:2 * dx + 5]).sum()
                if ws > 0:
                    ref = (k[:, None, None] * k[None, :, None] * num[2 * dy:2 * dy + 5, 2 * dx:2 * dx + 5]).sum(
                        axis=(0, 1)) / ws
                    assert np.allclose(got[dy, dx], ref, atol=1e-12) and gs[dy, dx] == 1
                else:
                    assert gs[dy, dx] == 0 and not got[dy, dx].any()
        # uniform preserved under arbitrary support
        u = np.full((h, w), 0.37)
        gu, _ = oracle64.reduce_support(u, s)
        assert np.allclose(gu[gs > 0], 0.37, atol=1e-14)


def _expand_nowrap_brute(src, dw, dh):
    """Truncated 5-tap expand renormalised by the in-bounds weight, with the divisors the
    reference uses when the source level is one pixel wide or tall
    (src/fuse/pyramid.zig:1231,1246,1248 for height 1; :1308-1309 for width 1)."""
    sh, sw = src.shape[:2]
    out = np.zeros((dh, dw, 3))

    def tap(y, x):
        return src[y, x] if (0 <= y < sh and 0 <= x < sw) else np.zeros(3)

    for y in range(dh):
        j, b = divmod(y, 2)
        for x in range(dw):
            i, a = divmod(x, 2)
            if a == 0:
                hk = [(i - 1, 1.0), (i, 6.0), (i + 1, 1.0)]
            else:
                hk = [(i, 4.0), (i + 1, 4.0)]
            if b == 0:
                vk = [(j - 1, 1.0), (j, 6.0), (j + 1, 1.0)]
            else:
                vk = [(j, 1.0), (j + 1, 1.0)]
            num = sum(vw * hw * tap(yy, xx) for (yy, vw) in vk for (xx, hw) in hk)
            hs = sum(hw for (xx, hw) in hk if 0 <= xx < sw)
            vs = sum(vw for (yy, vw) in vk if 0 <= yy < sh)
            div = hs * vs
            if sw == 1 and sh > 1 and j <= sh - 2:
                if a == 0 and b == 0:
                    div = 49.0 if j == 0 else 56.0
                elif a == 0 and b == 1:
                    div = 14.0
            if sh == 1 and sw > 1 and y == 0 and i == sw - 1:
                div = 49.0 if a == 0 else 28.0
            out[y, x] = num / div
    return out


@pytest.mark.parametrize("sw,sh", [(5, 4), (4, 5), (1, 1), (1, 4), (4, 1), (2, 1), (1, 2), (3, 3), (7, 2)])
def test_expand_nowrap_brute_force(oracle64, sw, sh):
    rng = np.random.default_rng(7)
    src = rng.random((sh, sw, 3))
    for dw in (2 * sw - 1, 2 * sw):
        for dh in (2 * sh - 1, 2 * sh):
            if dw < 1 or dh < 1:
                continue
            got = oracle64.expand(src, dw, dh)
            ref = _expand_nowrap_brute(src, dw, dh)
            assert np.allclose(got, ref, rtol=0, atol=1e-12), (sw, sh, dw, dh)


def test_expand_clamp_brute_force(oracle64):
    """>512 px: clamp-to-edge separable expand; odd output rows repeat source row y/2
    (src/fuse/pyramid.zig:1127-1128 fetch the same row twice)."""
    rng = np.random.default_rng(8)
    sw, sh = 300, 5
    src = rng.random((sh, sw, 3))
    dw, dh = 2 * sw - 1, 2 * sh

    def hrow(r):
        row = np.zeros((dw, 3))
        for x in range(dw):
            i = x // 2
            c = lambda q: min(max(q, 0), sw - 1)
            row[x] = (src[r, c(i - 1)] + 6 * src[r, c(i)] + src[r, c(i + 1)]) / 8 if x % 2 == 0 else \
                (src[r, c(i)] + src[r, c(i + 1)]) / 2
        return row

    rows = [hrow(r) for r in range(sh)]
    ref = np.zeros((dh, dw, 3))
    for y in range(dh):
        j = y // 2
        c = lambda q: min(max(q, 0), sh - 1)
        ref[y] = (rows[c(j - 1)] + 6 * rows[c(j)] + rows[c(j + 1)]) / 8 if y % 2 == 0 else rows[c(j)]
    got = oracle64.expand(src, dw, dh)
    assert np.allclose(got, ref, atol=1e-12)


def test_reduce_clamped_brute_force(oracle64):
    rng = np.random.default_rng(9)
    src = rng.random((7, 10, 3))
    got = oracle64.reduce_clamped(src)
    k = np.array([1, 4, 6, 4, 1]) / 16.0
    ref = np.zeros_like(got)
    for dy in range(got.shape[0]):
        for dx in range(got.shape[1]):
            acc = 0
            for ky in range(5):
                for kx in range(5):
                    yy = min(max(2 * dy + ky - 2, 0), 6)
                    xx = min(max(2 * dx + kx - 2, 0), 9)
                    acc = acc + k[ky] * k[kx] * src[yy, xx]
            ref[dy, dx] = acc
    assert np.allclose(got, ref, atol=1e-12)


# --------------------------------------------------------------------------
# whole-stack invariants
# --------------------------------------------------------------------------

def _frames(n, h, w, seed, alpha=True, dtype=np.uint8):
    rng = np.random.default_rng(seed)
    maxv = 255 if dtype == np.uint8 else 65535
    fr = []
    for _ in range(n):
        f = rng.integers(0, maxv + 1, (h, w, 4 if alpha else 3)).astype(dtype)
        if alpha:
            f[..., 3] = maxv
            f[:2, :, 3] = 0
            f[:, -3:, 3] = rng.integers(0, maxv, (h, 3))
        fr.append(f)
    return fr


@pytest.mark.parametrize("h,w", [(33, 47), (600, 70)])
def test_perfect_reconstruction(oracle, h, w):
    """collapse(laplacian(img)) == img: hybrid merge with zero mix returns the soft image."""
    rng = np.random.default_rng(10)
    img = rng.integers(0, 256, (h, w, 3)).astype(np.uint8)
    out = oracle.merge_hybrid(img, np.zeros_like(img), 0.0)
    assert np.array_equal(out, img)


@pytest.mark.parametrize("n", [1, 3])
def test_identical_frames_reproduce_input(oracle, n):
    f = _frames(1, 45, 61, 11, alpha=False)[0]
    out = oracle.fuse_frames([f] * n, extra=0)
    assert np.array_equal(out, f)


def test_masks_sum_to_one(oracle64):
    frames = _frames(4, 30, 41, 12)
    _, dbg = oracle64.fuse_frames(frames, debug=True)
    w = dbg["weights"]
    S = dbg["norm_sum"]
    assert np.allclose(w.sum(axis=0), S, rtol=1e-12)
    m = np.where(S > 1e-12, w / np.where(S > 0, S, 1), 0.25)
    assert np.allclose(m.sum(axis=0), 1.0, atol=1e-12)


def test_hybrid_zero_sharpness_is_pyramid(oracle):
    frames = _frames(3, 40, 52, 13)
    pyr = oracle.fuse_frames(frames, method=0)
    hyb = oracle.fuse_frames(frames, method=1, sharpness=0.0)
    assert np.array_equal(pyr, hyb)


def test_f32_oracle_within_tolerance_of_f64(oracle, oracle64):
    """The reference's f32 storage stays within the stated tolerance of the f64 restatement."""
    import synth
    spec = synth.config_spec(0, frames=4, width=160, height=120)
    st = synth.Stack(spec)
    src = [st.frame_numpy(k) for k in range(4)]
    from oracle import make_pose as mp
    poses = [mp(**p) for p in st.poses]
    roi = oracle.common_overlap_roi([(160, 120)] * 4, poses)
    a = oracle.stack(src, poses, roi).astype(int)
    b = oracle64.stack(src, poses, roi).astype(int)
    d = np.abs(a - b)
    assert d.max() <= 1 and d.mean() < 0.05


# ---- single-pass modes (hardmask-contrast, softmask-contrast) ----

def test_blur_keeps_uniform_weights(oracle):
    # src/fuse/masks.zig:209-226
    out = oracle.blur_five_tap(np.ones((3, 3), np.float32))
    assert np.allclose(out, 1.0, atol=1e-6)


@pytest.mark.parametrize("h,w", [(1, 1), (2, 7), (9, 5), (17, 23)])
def test_blur_brute_force(oracle, h, w):
    """Edge padding and whole-array slices instead of per-index clamping; same f32 sum order."""
    g = np.random.default_rng(h * 100 + w)
    x = g.random((h, w), dtype=np.float32) * np.float32(3)
    p = np.pad(x, ((0, 0), (2, 2)), mode="edge")
    hb = ((((p[:, 0:w] + np.float32(4) * p[:, 1:w + 1]) + np.float32(6) * p[:, 2:w + 2]) +
           np.float32(4) * p[:, 3:w + 3]) + p[:, 4:w + 4]) / np.float32(16)
    q = np.pad(hb, ((2, 2), (0, 0)), mode="edge")
    vb = ((((q[0:h] + np.float32(4) * q[1:h + 1]) + np.float32(6) * q[2:h + 2]) +
           np.float32(4) * q[3:h + 3]) + q[4:h + 4]) / np.float32(16)
    assert np.array_equal(oracle.blur_five_tap(x), vb)


def test_soft_blend_known_answer(oracle):
    # src/fuse/blend.zig:403-436: weights 1 and 3 over (10,20,30) and (110,120,130)
    ws = np.zeros((1, 1), np.float32)
    ps = np.zeros((1, 1, 3), np.float32)
    oracle.soft_accumulate(np.array([[[10, 20, 30]]], np.uint8), np.array([[1.0]], np.float32), ws, ps, extra=0)
    oracle.soft_accumulate(np.array([[[110, 120, 130]]], np.uint8), np.array([[3.0]], np.float32), ws, ps, extra=0)
    assert oracle.soft_finalize(ws, ps).ravel().tolist() == [85, 95, 105]


def test_soft_blend_skips_nonpositive_and_empty(oracle):
    ws = np.zeros((1, 2), np.float32)
    ps = np.zeros((1, 2, 3), np.float32)
    img = np.array([[[200, 100, 50, 255], [7, 8, 9, 255]]], np.uint8)
    oracle.soft_accumulate(img, np.array([[-1.0, 0.0]], np.float32), ws, ps)
    assert ws.tolist() == [[0.0, 0.0]]
    assert oracle.soft_finalize(ws, ps).ravel().tolist() == [0] * 6


def test_single_pass_modes_on_identical_frames(oracle):
    """N identical frames: every method returns the frame (hardmask: frame 0 wins every tie;
    softmask: the weighted mean of equal colours, wherever some weight is positive)."""
    f = _frames(1, 30, 41, 5)[0]
    frames = [f, f.copy(), f.copy()]
    from oracle import METHOD_HARDMASK, METHOD_SOFTMASK
    hard = oracle.fuse_frames(frames, method=METHOD_HARDMASK)
    assert np.array_equal(hard, f[..., :3])
    soft = oracle.fuse_frames(frames, method=METHOD_SOFTMASK)
    keep = soft.sum(axis=2) > 0
    assert keep.mean() > 0.5
    assert np.abs(soft.astype(int) - f[..., :3].astype(int))[keep].max() <= 1


def test_softmask_matches_numpy_restatement(oracle64):
    """Softmask end to end against a numpy composition of the pinned pieces (f64 build)."""
    from oracle import METHOD_SOFTMASK
    frames = _frames(3, 24, 19, 11)
    out = oracle64.fuse_frames(frames, method=METHOD_SOFTMASK)
    h, w = frames[0].shape[:2]
    ws = np.zeros((h, w)); ps = np.zeros((h, w, 3))
    for f in frames:
        wk = oracle64.contrast(oracle64.gray(f), oracle64.binary_support(f), 5)
        wk = wk * (f[..., 3] / 255.0)
        wk = oracle64.blur_five_tap(wk)
        pos = wk > 0
        ws[pos] += wk[pos]
        ps[pos] += wk[pos][:, None] * f[..., :3][pos]
    ref = np.where(ws[..., None] > 1e-12, np.clip(np.floor(ps / np.where(ws > 0, ws, 1)[..., None] + 0.5), 0, 255), 0)
    assert np.abs(out.astype(int) - ref.astype(int)).max() <= 1


# ---- hybrid mix and masked laplacian blend, restated from the pinned primitives ----
# reduce_clamped, expand (both regimes), reduce_support, contrast, gray, normalise and the
# collapse's quantisation are each pinned above by brute force or the reference's known answers;
# these compose them in numpy, as the reference composes them, and check the oracle's own
# composition against that.  Each test also shows that the mistake it guards against changes the
# result (so the comparison is not vacuous).

def _lap_pyramid(o, img, levels):
    g = [np.asarray(img, np.float64)]
    for _ in range(1, levels):
        g.append(o.reduce_clamped(g[-1]))
    lap = [g[l] - o.expand(g[l + 1], g[l].shape[1], g[l].shape[0]) for l in range(levels - 1)]
    return lap + [g[-1]]


def _collapse_quantize(o, levels, hi):
    c = [l.copy() for l in levels]
    for l in range(len(c) - 1, 0, -1):
        c[l - 1] = c[l - 1] + o.expand(c[l], c[l - 1].shape[1], c[l - 1].shape[0])
    return np.floor(np.clip(c[0] + 0.5, 0.0, hi)).astype(np.int64)   # (int)clamp(v + 0.5), v + 0.5 >= 0


def _hybrid_restated(o, soft, hard, mix, start=2, count=4, hi=255, levels=None):
    """mergeHybridImages src/fuse/pyramid.zig:271-301: laplacian pyramids of both images (reduceRgb
    clamped, :418-487), L = soft + (hard - soft) * mix on levels [start, start + count), collapse."""
    h, w = soft.shape[:2]
    n = levels or len(level_dims(w, h))
    ls, lh = _lap_pyramid(o, soft, n), _lap_pyramid(o, hard, n)
    m = float(np.float32(mix))
    hs = min(start, n)
    he = min(hs + count, n)
    mixed = [ls[l] + (lh[l] - ls[l]) * m if hs <= l < he else ls[l] for l in range(n)]
    return _collapse_quantize(o, mixed, hi)


@pytest.mark.parametrize("mix", [0.35, 1.0, 0.6])
@pytest.mark.parametrize("h,w,dtype", [(110, 150, np.uint8), (37, 70, np.uint16), (700, 530, np.uint8)])
def test_hybrid_merge_matches_numpy_restatement(oracle64, mix, h, w, dtype):
    g = np.random.default_rng(h * 7 + w)
    hi = np.iinfo(dtype).max
    soft = g.integers(0, hi + 1, (h, w, 3)).astype(dtype)
    hard = g.integers(0, hi + 1, (h, w, 3)).astype(dtype)
    got = oracle64.merge_hybrid(soft, hard, mix).astype(np.int64)
    ref = _hybrid_restated(oracle64, soft, hard, mix, hi=hi)
    assert np.array_equal(got, ref)
    # the slips this catches: operands swapped, the range shifted by one level either way
    for wrong in (_hybrid_restated(oracle64, hard, soft, mix, hi=hi),
                  _hybrid_restated(oracle64, soft, hard, mix, start=1, hi=hi),
                  _hybrid_restated(oracle64, soft, hard, mix, start=3, hi=hi),
                  _hybrid_restated(oracle64, soft, hard, mix, count=3, hi=hi)):
        assert np.abs(wrong - got).max() > 1


def test_hybrid_merge_sharpness_one_takes_the_hard_band(oracle64):
    """At mix 1 levels [2, 6) are the hard image's own laplacian: a hard image that differs from the
    soft one only in its band-pass content moves the output, one that differs only in its coarsest
    levels (a constant offset) does not."""
    g = np.random.default_rng(3)
    soft = g.integers(40, 200, (96, 128, 3)).astype(np.uint8)
    const = (soft.astype(int) + 17).astype(np.uint8)
    out = oracle64.merge_hybrid(soft, const, 1.0)
    assert np.abs(out.astype(int) - soft.astype(int)).max() <= 1
    out_soft = oracle64.merge_hybrid(soft, g.integers(0, 256, soft.shape).astype(np.uint8), 1.0)
    assert np.abs(out_soft.astype(int) - soft.astype(int)).max() > 8


@pytest.mark.parametrize("levels", [3, 6, 7])
def test_hybrid_merge_capped_levels_restated(oracle64, levels):
    """With the pyramid capped the hard range is clipped to the levels that exist (:286-287)."""
    g = np.random.default_rng(levels)
    soft = g.integers(0, 256, (90, 120, 3)).astype(np.uint8)
    hard = g.integers(0, 256, (90, 120, 3)).astype(np.uint8)
    got = oracle64.merge_hybrid(soft, hard, 0.35, max_levels=levels).astype(np.int64)
    assert np.array_equal(got, _hybrid_restated(oracle64, soft, hard, 0.35, levels=levels))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_hybrid_stack_is_merge_of_pyramid_and_hardmask(oracle, dtype):
    """src/stack/pipeline.zig:177-236: the hybrid's soft image is the pyramid-contrast result and its
    hard image the hardmask-contrast result of the same frames; the output is their merge."""
    from oracle import METHOD_PYRAMID, METHOD_HYBRID, METHOD_HARDMASK
    frames = _frames(4, 52, 71, 21, dtype=dtype)
    soft = oracle.fuse_frames(frames, method=METHOD_PYRAMID)
    hard = oracle.fuse_frames(frames, method=METHOD_HARDMASK)
    for s in (0.35, 1.0):
        hyb, dbg = oracle.fuse_frames(frames, method=METHOD_HYBRID, sharpness=s, debug=True)
        assert np.array_equal(dbg["soft"], soft) and np.array_equal(dbg["hard"], hard)
        assert np.array_equal(hyb, oracle.merge_hybrid(soft, hard, s))


def _alpha_frames(n, h, w, seed, dtype=np.uint8):
    """Aligned frames whose alpha varies per frame (partial values, and holes in different places),
    so that each frame's own support differs from the union support."""
    g = np.random.default_rng(seed)
    hi = np.iinfo(dtype).max
    out = []
    for k in range(n):
        f = g.integers(0, hi + 1, (h, w, 4)).astype(dtype)
        a = np.full((h, w), hi, np.int64)
        a[: 3 + 2 * k, :] = 0                                   # top band, a different height per frame
        a[:, w - 4 - k:] = g.integers(0, hi + 1, (h, 4 + k))    # right band of partial alpha
        a[h // 2 - 5: h // 2 + 5, 7 * k: 7 * k + 9] = 0        # a hole at a different place per frame
        f[..., 3] = a
        out.append(f)
    return out


def _pyramid_blend_restated(o, frames, window=5, own_mask_support=False):
    """accumulateImageWithWorkspace src/fuse/pyramid.zig:190-204: per frame, the normalised mask's
    gaussian pyramid reduced with the UNION support at level 0 (:202, :384-416), the image's
    gaussian pyramid reduced with the frame's OWN alpha > 0 support (:489-542), the laplacian times
    the mask added level by level (:1506-1554); then the collapse (:223-269)."""
    n = len(frames)
    h, w = frames[0].shape[:2]
    hi = np.iinfo(frames[0].dtype).max
    dims = level_dims(w, h)
    L = len(dims)
    wts = []
    for f in frames:
        wk = o.contrast(o.gray(f), o.binary_support(f), window)
        wts.append(wk * (f[..., 3].astype(np.float64) / hi))
    S = np.zeros((h, w))
    for wk in wts:
        S = S + wk
    U = np.zeros((h, w))
    for f in frames:
        U = np.maximum(U, (f[..., 3] > 0).astype(np.float64))
    acc = [np.zeros((lh, lw, 3)) for (lw, lh) in dims]
    for f, wk in zip(frames, wts):
        m = [np.where(S > 1e-12, wk / np.where(S > 1e-12, S, 1.0), 1.0 / n)]
        ms = [(f[..., 3] > 0).astype(np.float64) if own_mask_support else U]
        for _ in range(1, L):
            a, b = o.reduce_support(m[-1], ms[-1])
            m.append(a)
            ms.append(b)
        g = [f[..., :3].astype(np.float64)]
        gs = [(f[..., 3] > 0).astype(np.float64)]
        for _ in range(1, L):
            a, b = o.reduce_support(g[-1], gs[-1])
            g.append(a)
            gs.append(b)
        for l in range(L - 1):
            lap = g[l] - o.expand(g[l + 1], dims[l][0], dims[l][1])
            acc[l] = acc[l] + lap * m[l][..., None]
        acc[L - 1] = acc[L - 1] + g[L - 1] * m[L - 1][..., None]
    return _collapse_quantize(o, acc, hi)


@pytest.mark.parametrize("h,w,dtype", [(48, 64, np.uint8), (61, 45, np.uint16), (600, 560, np.uint8)])
def test_masked_pyramid_blend_matches_numpy_restatement(oracle64, h, w, dtype):
    frames = _alpha_frames(3, h, w, h + w, dtype)
    got = oracle64.fuse_frames(frames, extra=1).astype(np.int64)
    assert np.array_equal(got, _pyramid_blend_restated(oracle64, frames))
    # the slip this catches: the frame's own support in place of the union for the mask pyramid
    wrong = _pyramid_blend_restated(oracle64, frames, own_mask_support=True)
    assert np.abs(wrong - got).max() > 1
