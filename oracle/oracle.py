"""ctypes front end of the CPU oracle (oracle/focus_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, to check results.  It restates the reference algorithm;
it is never the thing measured or shipped.

`Oracle(fp64=False)` wraps liboracle.so (the reference's f32/f64 mix);
`Oracle(fp64=True)` wraps liboracle64.so (all f32 storage promoted to f64).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

POSE_FIELDS = (
    "yaw", "pitch", "roll", "hfov_delta", "trans_x", "trans_y", "trans_z",
    "translation_plane_yaw", "translation_plane_pitch", "radial_a", "radial_b", "radial_c",
    "center_shift_x", "center_shift_y", "base_hfov_degrees",
)



# OR_METHOD_* of focus_oracle.h
METHOD_PYRAMID, METHOD_HYBRID, METHOD_HARDMASK, METHOD_SOFTMASK = 0, 1, 2, 3

class Pose(C.Structure):
    _fields_ = [(f, C.c_double) for f in POSE_FIELDS]


def make_pose(**kw) -> Pose:
    p = Pose()
    p.base_hfov_degrees = 50.0
    for k, v in kw.items():
        setattr(p, k, float(v))
    return p


class InvCache(C.Structure):
    _fields_ = [
        ("basic_rectilinear", C.c_int), ("has_translation", C.c_int),
        ("src_center_x", C.c_double), ("src_center_y", C.c_double),
        ("dest_center_x", C.c_double), ("dest_center_y", C.c_double),
        ("pano_distance", C.c_double), ("half_turn", C.c_double), ("resize_scale", C.c_double),
        ("dest_focal", C.c_double), ("image_focal", C.c_double),
        ("pitch_roll", C.c_double * 9), ("world_to_local", C.c_double * 9),
        ("yaw_turn", C.c_double), ("yaw", C.c_double),
        ("radial_a", C.c_double), ("radial_b", C.c_double), ("radial_c", C.c_double),
        ("center_shift_x", C.c_double), ("center_shift_y", C.c_double), ("radial_inv_norm", C.c_double),
        ("trans_x", C.c_double), ("trans_y", C.c_double), ("trans_z", C.c_double),
        ("plane_normal", C.c_double * 3),
    ]


class FuseOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("window", C.c_uint32), ("sharpness", C.c_float),
                ("max_levels", C.c_uint32)]


class FuseDebug(C.Structure):
    _fields_ = [("norm_sum", C.c_void_p), ("union_support", C.c_void_p), ("weights", C.c_void_p),
                ("acc_levels", C.c_void_p), ("mask_levels", C.c_void_p), ("gauss_levels", C.c_void_p),
                ("frame_index", C.c_int), ("soft", C.c_void_p), ("hard", C.c_void_p)]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def level_dims(w: int, h: int, max_levels: int = 0):
    dims = [(w, h)]
    while w > 1 or h > 1:
        w = (w + 1) // 2 if w > 1 else 1
        h = (h + 1) // 2 if h > 1 else 1
        dims.append((w, h))
    if max_levels:
        dims = dims[:max_levels]
    return dims


def split_levels(flat: np.ndarray, dims, channels: int):
    out, o = [], 0
    for (w, h) in dims:
        n = w * h * channels
        out.append(flat[o:o + n].reshape(h, w, channels) if channels > 1 else flat[o:o + n].reshape(h, w))
        o += n
    return out


class Oracle:
    def __init__(self, fp64: bool = False):
        name = "liboracle64.so" if fp64 else "liboracle.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `python oracle/build.py`")
        self.lib = C.CDLL(path)
        self.real = np.float64 if fp64 else np.float32
        L = self.lib
        vp, u32, i32, d = C.c_void_p, C.c_uint32, C.c_int, C.c_double
        L.or_init_inverse_cache.argtypes = [C.POINTER(Pose), u32, u32, C.POINTER(InvCache)]
        L.or_inverse_transform_cached.argtypes = [C.POINTER(InvCache), d, d, C.POINTER(d), C.POINTER(d)]
        L.or_inverse_legacy_chain.argtypes = [C.POINTER(InvCache), d, d, C.POINTER(d), C.POINTER(d)]
        L.or_transform_point.argtypes = [C.POINTER(Pose), d, d, u32, u32, C.POINTER(d), C.POINTER(d)]
        L.or_common_overlap_roi.argtypes = [i32, vp, vp, vp, vp, vp]
        L.or_common_overlap_roi.restype = i32
        L.or_remap_u8.argtypes = [vp, u32, u32, i32, C.POINTER(Pose), vp, vp]
        L.or_remap_u16.argtypes = [vp, u32, u32, i32, C.POINTER(Pose), vp, vp]
        for fn in ("or_gray", "or_binary_support", "or_union_support", "or_apply_support", "or_fill_rgb_base"):
            getattr(L, fn).argtypes = [vp, i32, u32, u32, i32, i32, vp]
        L.or_contrast.argtypes = [vp, vp, u32, u32, u32, vp]
        L.or_sample_std.argtypes = [d, d, u32]
        L.or_sample_std.restype = d
        L.or_normalize.argtypes = [vp, vp, u32, u32, vp]
        L.or_level_count.argtypes = [u32, u32]
        L.or_level_count.restype = u32
        L.or_reduce_scalar_support.argtypes = [u32, u32, vp, vp, u32, u32, vp, vp]
        L.or_reduce_rgb_support.argtypes = [u32, u32, vp, vp, u32, u32, vp, vp]
        L.or_reduce_rgb_clamped.argtypes = [u32, u32, vp, u32, u32, vp]
        L.or_expand_rgb.argtypes = [u32, u32, u32, u32, vp, vp]
        L.or_fuse_frames.argtypes = [i32, vp, u32, u32, i32, i32, C.POINTER(FuseOpts), vp, C.POINTER(FuseDebug)]
        L.or_fuse_frames.restype = i32
        L.or_stack.argtypes = [i32, vp, u32, u32, i32, vp, vp, C.POINTER(FuseOpts), vp, C.POINTER(FuseDebug)]
        L.or_stack.restype = i32
        L.or_update_winners.argtypes = [vp, i32, u32, u32, i32, i32, vp, vp, vp]
        L.or_blur_five_tap.argtypes = [u32, u32, vp, vp, vp]
        L.or_soft_accumulate.argtypes = [vp, i32, u32, u32, i32, i32, vp, vp, vp]
        L.or_soft_finalize.argtypes = [vp, vp, u32, u32, i32, vp]
        L.or_merge_hybrid.argtypes = [vp, vp, i32, u32, u32, u32, u32, C.c_float, u32, vp]
        L.or_collapse.argtypes = [u32, u32, u32, vp, i32, vp]
        L.or_set_threads.argtypes = [i32]
        L.or_get_threads.restype = i32

    # ---- threads of the parallel loops (the same bits at any count) ----
    def set_threads(self, n: int) -> None:
        self.lib.or_set_threads(int(n))

    def threads(self) -> int:
        return int(self.lib.or_get_threads())

    # ---- pose math ----
    def inverse_cache(self, pose: Pose, w: int, h: int) -> InvCache:
        c = InvCache()
        self.lib.or_init_inverse_cache(C.byref(pose), w, h, C.byref(c))
        return c

    def inverse_point(self, pose: Pose, w: int, h: int, x: float, y: float, legacy_chain: bool = False):
        c = self.inverse_cache(pose, w, h)
        sx, sy = C.c_double(), C.c_double()
        fn = self.lib.or_inverse_legacy_chain if legacy_chain else self.lib.or_inverse_transform_cached
        fn(C.byref(c), x, y, C.byref(sx), C.byref(sy))
        return sx.value, sy.value

    def transform_point(self, pose: Pose, w: int, h: int, x: float, y: float):
        ox, oy = C.c_double(), C.c_double()
        self.lib.or_transform_point(C.byref(pose), x, y, w, h, C.byref(ox), C.byref(oy))
        return ox.value, oy.value

    def common_overlap_roi(self, sizes, poses, active=None):
        n = len(poses)
        ws = np.array([s[0] for s in sizes], np.uint32)
        hs = np.array([s[1] for s in sizes], np.uint32)
        arr = (Pose * n)(*poses)
        act = None if active is None else np.asarray(active, np.uint8)
        roi = np.zeros(4, np.int32)
        ok = self.lib.or_common_overlap_roi(n, None if act is None else _ptr(act), _ptr(ws), _ptr(hs),
                                            C.cast(arr, C.c_void_p), _ptr(roi))
        return tuple(int(v) for v in roi) if ok else None

    # ---- remap ----
    def remap(self, src: np.ndarray, pose: Pose, roi=None) -> np.ndarray:
        src = np.ascontiguousarray(src)
        h, w = src.shape[:2]
        ch = 1 if src.ndim == 2 else src.shape[2]
        r = (0, 0, w, h) if roi is None else tuple(roi)
        out = np.zeros((r[3] - r[1], r[2] - r[0], ch + 1), src.dtype)
        roi_a = np.array(r, np.int32)
        fn = self.lib.or_remap_u8 if src.dtype == np.uint8 else self.lib.or_remap_u16
        fn(_ptr(src), w, h, ch, C.byref(pose), _ptr(roi_a), _ptr(out))
        return out

    # ---- building blocks ----
    def _img_call(self, fn, img, extra, out):
        img = np.ascontiguousarray(img)
        h, w, cc = img.shape
        color = cc - extra
        getattr(self.lib, fn)(_ptr(img), img.dtype.itemsize, w, h, color, extra, _ptr(out))
        return out

    def gray(self, img, extra=1):
        h, w = img.shape[:2]
        return self._img_call("or_gray", img, extra, np.zeros((h, w), self.real))

    def binary_support(self, img, extra=1):
        h, w = img.shape[:2]
        return self._img_call("or_binary_support", img, extra, np.zeros((h, w), self.real))

    def contrast(self, gray, support, window=5):
        gray = np.ascontiguousarray(gray, self.real)
        support = np.ascontiguousarray(support, self.real)
        h, w = gray.shape
        out = np.zeros((h, w), self.real)
        self.lib.or_contrast(_ptr(gray), _ptr(support), w, h, window, _ptr(out))
        return out

    def sample_std(self, s, q, n):
        return self.lib.or_sample_std(s, q, n)

    def normalize(self, weights, norm, total):
        weights = np.ascontiguousarray(weights, self.real)
        norm = np.ascontiguousarray(norm, self.real)
        out = np.zeros_like(weights)
        self.lib.or_normalize(_ptr(weights), _ptr(norm), weights.size, total, _ptr(out))
        return out

    def level_count(self, w, h):
        return self.lib.or_level_count(w, h)

    def reduce_support(self, src, support):
        src = np.ascontiguousarray(src, self.real)
        support = np.ascontiguousarray(support, self.real)
        sh, sw = support.shape
        dw, dh = (sw + 1) // 2 if sw > 1 else 1, (sh + 1) // 2 if sh > 1 else 1
        dsup = np.zeros((dh, dw), self.real)
        if src.ndim == 3:
            dst = np.zeros((dh, dw, 3), self.real)
            self.lib.or_reduce_rgb_support(sw, sh, _ptr(src), _ptr(support), dw, dh, _ptr(dst), _ptr(dsup))
        else:
            dst = np.zeros((dh, dw), self.real)
            self.lib.or_reduce_scalar_support(sw, sh, _ptr(src), _ptr(support), dw, dh, _ptr(dst), _ptr(dsup))
        return dst, dsup

    def reduce_clamped(self, src):
        src = np.ascontiguousarray(src, self.real)
        sh, sw = src.shape[:2]
        dw, dh = (sw + 1) // 2 if sw > 1 else 1, (sh + 1) // 2 if sh > 1 else 1
        dst = np.zeros((dh, dw, 3), self.real)
        self.lib.or_reduce_rgb_clamped(sw, sh, _ptr(src), dw, dh, _ptr(dst))
        return dst

    def expand(self, src, dw, dh):
        src = np.ascontiguousarray(src, self.real)
        sh, sw = src.shape[:2]
        dst = np.zeros((dh, dw, 3), self.real)
        self.lib.or_expand_rgb(dw, dh, sw, sh, _ptr(src), _ptr(dst))
        return dst

    def collapse(self, levels, dtype=np.uint8):
        h, w = levels[0].shape[:2]
        flat = np.concatenate([np.ascontiguousarray(l, self.real).ravel() for l in levels])
        out = np.zeros((h, w, 3), dtype)
        self.lib.or_collapse(w, h, len(levels), _ptr(flat), np.dtype(dtype).itemsize, _ptr(out))
        return out

    def update_winners(self, img, raw_weights, best, out, extra=1):
        img = np.ascontiguousarray(img)
        h, w, cc = img.shape
        raw = np.ascontiguousarray(raw_weights, self.real)
        self.lib.or_update_winners(_ptr(img), img.dtype.itemsize, w, h, cc - extra, extra, _ptr(raw), _ptr(best),
                                   _ptr(out))

    def blur_five_tap(self, img):
        src = np.ascontiguousarray(img, self.real)
        h, w = src.shape
        scratch = np.zeros_like(src)
        out = np.zeros_like(src)
        self.lib.or_blur_five_tap(w, h, _ptr(src), _ptr(scratch), _ptr(out))
        return out

    def soft_accumulate(self, img, weights, weight_sums, pixel_sums, extra=1):
        img = np.ascontiguousarray(img)
        h, w, cc = img.shape
        wt = np.ascontiguousarray(weights, self.real)
        self.lib.or_soft_accumulate(_ptr(img), img.dtype.itemsize, w, h, cc - extra, extra, _ptr(wt),
                                    _ptr(weight_sums), _ptr(pixel_sums))

    def soft_finalize(self, weight_sums, pixel_sums, dtype=np.uint8):
        h, w = weight_sums.shape
        out = np.zeros((h, w, 3), dtype)
        self.lib.or_soft_finalize(_ptr(np.ascontiguousarray(weight_sums, self.real)),
                                  _ptr(np.ascontiguousarray(pixel_sums, self.real)), w, h, np.dtype(dtype).itemsize,
                                  _ptr(out))
        return out

    def merge_hybrid(self, soft, hard, mix, start=2, count=4, max_levels=0):
        soft = np.ascontiguousarray(soft)
        hard = np.ascontiguousarray(hard)
        h, w = soft.shape[:2]
        out = np.zeros_like(soft)
        self.lib.or_merge_hybrid(_ptr(soft), _ptr(hard), soft.dtype.itemsize, w, h, start, count, mix,
                                 max_levels, _ptr(out))
        return out

    # ---- whole stack ----
    def _debug(self, n, w, h, dims, frame_index, method):
        cnt = w * h
        tot = sum(a * b for a, b in dims)
        d = {
            "norm_sum": np.zeros(cnt, self.real), "union_support": np.zeros(cnt, self.real),
            "weights": np.zeros(cnt * n, self.real), "acc_levels": np.zeros(tot * 3, self.real),
            "mask_levels": np.zeros(tot, self.real), "gauss_levels": np.zeros(tot * 3, self.real),
        }
        dbg = FuseDebug()
        for k, v in d.items():
            setattr(dbg, k, v.ctypes.data)
        dbg.frame_index = frame_index
        return dbg, d

    def _finish_debug(self, d, n, w, h, dims, dtype, dbg_struct, soft, hard):
        out = {
            "norm_sum": d["norm_sum"].reshape(h, w), "union_support": d["union_support"].reshape(h, w),
            "weights": d["weights"].reshape(n, h, w),
            "acc_levels": split_levels(d["acc_levels"], dims, 3),
            "mask_levels": split_levels(d["mask_levels"], dims, 1),
            "gauss_levels": split_levels(d["gauss_levels"], dims, 3),
        }
        if soft is not None:
            out["soft"], out["hard"] = soft, hard
        return out

    def fuse_frames(self, frames, extra=1, method=0, window=5, sharpness=0.35, max_levels=0, debug=False,
                    frame_index=0):
        frames = [np.ascontiguousarray(f) for f in frames]
        n = len(frames)
        h, w = frames[0].shape[:2]
        dtype = frames[0].dtype
        opts = FuseOpts(method, window, sharpness, max_levels)
        out = np.zeros((h, w, 3), dtype)
        arr = (C.c_void_p * n)(*[f.ctypes.data for f in frames])
        dims = level_dims(w, h, max_levels)
        dbg_p, d, soft, hard = None, None, None, None
        if debug:
            dbg, d = self._debug(n, w, h, dims, frame_index, method)
            if method == 1:
                soft, hard = np.zeros_like(out), np.zeros_like(out)
                dbg.soft, dbg.hard = soft.ctypes.data, hard.ctypes.data
            dbg_p = C.byref(dbg)
        rc = self.lib.or_fuse_frames(n, C.cast(arr, C.c_void_p), w, h, dtype.itemsize, extra, C.byref(opts),
                                     _ptr(out), dbg_p)
        if rc != 0:
            raise RuntimeError(f"or_fuse_frames failed: {rc}")
        if debug:
            return out, self._finish_debug(d, n, w, h, dims, dtype, None, soft, hard)
        return out

    def stack(self, sources, poses, roi=None, method=0, window=5, sharpness=0.35, max_levels=0, debug=False,
              frame_index=0):
        sources = [np.ascontiguousarray(s) for s in sources]
        n = len(sources)
        sh, sw = sources[0].shape[:2]
        r = (0, 0, sw, sh) if roi is None else tuple(roi)
        w, h = r[2] - r[0], r[3] - r[1]
        dtype = sources[0].dtype
        opts = FuseOpts(method, window, sharpness, max_levels)
        out = np.zeros((h, w, 3), dtype)
        arr = (C.c_void_p * n)(*[s.ctypes.data for s in sources])
        parr = (Pose * n)(*poses)
        roi_a = np.array(r, np.int32)
        dims = level_dims(w, h, max_levels)
        dbg_p, d, soft, hard = None, None, None, None
        if debug:
            dbg, d = self._debug(n, w, h, dims, frame_index, method)
            if method == 1:
                soft, hard = np.zeros_like(out), np.zeros_like(out)
                dbg.soft, dbg.hard = soft.ctypes.data, hard.ctypes.data
            dbg_p = C.byref(dbg)
        rc = self.lib.or_stack(n, C.cast(arr, C.c_void_p), sw, sh, dtype.itemsize, C.cast(parr, C.c_void_p),
                               _ptr(roi_a), C.byref(opts), _ptr(out), dbg_p)
        if rc != 0:
            raise RuntimeError(f"or_stack failed: {rc}")
        if debug:
            return out, self._finish_debug(d, n, w, h, dims, dtype, None, soft, hard)
        return out
