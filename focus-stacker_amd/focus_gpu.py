"""ctypes binding of libfocus_gpu.so (include/focus_gpu.h) for Python callers, tests and bench.

Mirrors the C ABI one to one and adds numpy conveniences.  There is no fallback: if the
HIP library is missing or no GPU is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfocus_gpu.so")

FG_OK = 0
ERRORS = {
    -1: "FG_ERR_INVALID_ARG", -2: "FG_ERR_OOM_HOST", -3: "FG_ERR_OOM_DEVICE", -4: "FG_ERR_SIZE_MISMATCH",
    -5: "FG_ERR_FORMAT_MISMATCH", -6: "FG_ERR_UNSUPPORTED", -7: "FG_ERR_DEVICE", -8: "FG_ERR_STATE",
}
SAMPLE_U8, SAMPLE_U16 = 0, 1
METHOD_HARDMASK, METHOD_SOFTMASK, METHOD_PYRAMID, METHOD_HYBRID = 0, 1, 2, 3
DEBUG = dict(norm_sum=0, union_support=1, weights=2, acc_level=3, mask_level=4, gauss_level=5, soft=6, hard=7)
# enum fg_phase order; pass2_level0 is the level-0 launch inside pass2
PHASES = ("ingest", "pass1", "pass2", "finish", "pass2_level0")

POSE_FIELDS = (
    "yaw", "pitch", "roll", "hfov_delta", "trans_x", "trans_y", "trans_z",
    "translation_plane_yaw", "translation_plane_pitch", "radial_a", "radial_b", "radial_c",
    "center_shift_x", "center_shift_y", "base_hfov_degrees",
)


class FgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Pose(C.Structure):
    _fields_ = [(f, C.c_double) for f in POSE_FIELDS]


class Rect(C.Structure):
    _fields_ = [("left", C.c_int32), ("top", C.c_int32), ("right", C.c_int32), ("bottom", C.c_int32)]


class ImageDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("color_channels", C.c_uint32),
                ("extra_channels", C.c_uint32), ("sample_type", C.c_uint32)]


class ContextOpts(C.Structure):
    _fields_ = [("device", C.c_int32), ("memory_fraction", C.c_float)]


class StackOpts(C.Structure):
    _fields_ = [("method", C.c_uint32), ("contrast_window_size", C.c_uint32), ("hybrid_sharpness", C.c_float),
                ("frame_count", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("sample_type", C.c_uint32), ("max_levels", C.c_uint32), ("memory_fraction", C.c_float),
                ("debug_frame", C.c_int32), ("stream", C.c_void_p)]


class ServerOpts(C.Structure):
    _fields_ = [("slots", C.c_int32), ("memory_fraction", C.c_float), ("pool_per_slot", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("frame_count", C.c_uint32), ("descs", C.c_void_p), ("frames", C.c_void_p), ("poses", C.c_void_p),
                ("roi", C.c_void_p), ("method", C.c_uint32), ("contrast_window_size", C.c_uint32),
                ("hybrid_sharpness", C.c_float), ("max_levels", C.c_uint32), ("out", C.c_void_p)]


class BatchOpts(C.Structure):
    _fields_ = [("memory_fraction", C.c_float), ("plan_cache", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("pass2_min_tiles", C.c_int32), ("batch_wide_level", C.c_int32),
                ("batch_pass1_blocks_per_cu", C.c_int32), ("batch_group_mib", C.c_int32),
                ("batch_streams", C.c_int32), ("batch_large_kpix", C.c_int32)]


class RequestStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("queue_ms", C.c_double), ("run_ms", C.c_double), ("total_ms", C.c_double)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE = 0, 1, 2


def pose(d: dict | None = None, **kw) -> Pose:
    p = Pose()
    p.base_hfov_degrees = 50.0
    for k, v in {**(d or {}), **kw}.items():
        setattr(p, k, float(v))
    return p


def rect(r) -> Rect | None:
    return None if r is None else Rect(*[int(v) for v in r])


def desc_for(a: np.ndarray, extra: int = 0, width: int | None = None, height: int | None = None) -> ImageDesc:
    ch = 1 if a.ndim == 2 else a.shape[-1]
    h, w = a.shape[:2] if a.ndim >= 2 else (height, width)
    if width is not None:
        w, h = width, height
    st = SAMPLE_U8 if a.dtype == np.uint8 else SAMPLE_U16
    return ImageDesc(w, h, ch - extra, extra, st)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python focus-stacker_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    vp, i32, u32 = C.c_void_p, C.c_int, C.c_uint32
    L.fg_last_error.restype = C.c_char_p
    L.fg_version.restype = C.c_char_p
    L.fg_context_create.argtypes = [C.POINTER(ContextOpts), C.POINTER(vp)]
    L.fg_context_destroy.argtypes = [vp]
    L.fg_context_destroy.restype = None
    L.fg_remap.argtypes = [vp, C.POINTER(ImageDesc), vp, C.POINTER(Pose), C.POINTER(Rect), vp]
    L.fg_remap_device.argtypes = [vp, C.POINTER(ImageDesc), vp, C.POINTER(Pose), C.POINTER(Rect), vp, vp]
    L.fg_common_overlap_roi.argtypes = [u32, vp, vp, vp, vp, C.POINTER(Rect)]
    L.fg_stack_create.argtypes = [vp, C.POINTER(StackOpts), C.POINTER(vp)]
    L.fg_stack_destroy.argtypes = [vp]
    L.fg_stack_destroy.restype = None
    for fn in ("fg_stack_add_frame", "fg_stack_add_frame_device", "fg_stack_pass2_frame",
               "fg_stack_pass2_frame_device"):
        getattr(L, fn).argtypes = [vp, C.POINTER(ImageDesc), vp, C.POINTER(Pose), C.POINTER(Rect)]
    L.fg_stack_needs_pass2_frames.argtypes = [vp]
    L.fg_stack_finish.argtypes = [vp, vp]
    L.fg_stack_finish_device.argtypes = [vp, vp]
    L.fg_stack_debug_readback.argtypes = [vp, i32, u32, vp, C.c_size_t]
    L.fg_stack_levels.argtypes = [vp, C.POINTER(u32)]
    L.fg_stack_level_size.argtypes = [vp, u32, C.POINTER(u32), C.POINTER(u32)]
    L.fg_stack_stream.argtypes = [vp]
    L.fg_stack_stream.restype = vp
    L.fg_stack_sync.argtypes = [vp]
    L.fg_stack_reset.argtypes = [vp]
    L.fg_stack_set_profiling.argtypes = [vp, i32]
    L.fg_stack_phase_times.argtypes = [vp, vp, vp]
    L.fg_fuse_stack.argtypes = [vp, u32, vp, vp, vp, C.POINTER(Rect), C.POINTER(StackOpts), vp]
    L.fg_fuse_stack_device_noalloc.argtypes = [vp, u32, vp, vp, vp, C.POINTER(Rect), C.POINTER(StackOpts), vp]
    L.fg_server_create.argtypes = [vp, C.POINTER(ServerOpts), C.POINTER(vp)]
    L.fg_server_destroy.argtypes = [vp]
    L.fg_server_destroy.restype = None
    L.fg_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_uint64)]
    L.fg_server_wait.argtypes = [vp, C.c_uint64, C.c_double, C.POINTER(RequestStatus)]
    L.fg_batch_create.argtypes = [vp, C.POINTER(BatchOpts), C.POINTER(vp)]
    L.fg_batch_destroy.argtypes = [vp]
    L.fg_batch_destroy.restype = None
    L.fg_batch_fuse.argtypes = [vp, u32, vp, vp]
    L.fg_get_tuning.argtypes = [C.POINTER(Tuning)]
    L.fg_set_tuning.argtypes = [C.POINTER(Tuning)]
    L.fg_device_alloc.argtypes = [vp, C.c_size_t]
    L.fg_device_alloc.restype = vp
    L.fg_device_free.argtypes = [vp, vp]
    L.fg_device_free.restype = None
    L.fg_copy_to_device.argtypes = [vp, vp, vp, C.c_size_t]
    L.fg_copy_to_host.argtypes = [vp, vp, vp, C.c_size_t]
    L.fg_host_alloc.argtypes = [C.c_size_t]
    L.fg_host_alloc.restype = vp
    L.fg_host_free.argtypes = [vp]
    L.fg_host_free.restype = None
    L.fg_write_tiff.argtypes = [C.c_char_p, C.POINTER(ImageDesc), vp]
    L.fg_write_aligned_tiff.argtypes = [vp, C.POINTER(ImageDesc), vp, C.POINTER(Pose), C.POINTER(Rect),
                                        C.c_char_p, u32]
    _lib = L
    return L


class PinnedArray:
    """A numpy array over page-locked host memory from fg_host_alloc (freed with close())."""

    def __init__(self, shape, dtype=np.uint8):
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        self.ptr = lib().fg_host_alloc(n)
        if not self.ptr:
            raise FgError(-2, lib().fg_last_error().decode())
        buf = (C.c_uint8 * n).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def close(self):
        if self.ptr:
            self.array = None
            lib().fg_host_free(self.ptr)
            self.ptr = None


def check(rc: int) -> int:
    if rc < 0:
        raise FgError(rc, lib().fg_last_error().decode())
    return rc


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    def __init__(self, device: int = 0, memory_fraction: float = 0.5):
        self.h = C.c_void_p()
        check(lib().fg_context_create(C.byref(ContextOpts(device, memory_fraction)), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().fg_context_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def alloc(self, nbytes: int) -> int:
        p = lib().fg_device_alloc(self.h, nbytes)
        if not p:
            raise FgError(-3, lib().fg_last_error().decode())
        return p

    def free(self, d_ptr: int):
        lib().fg_device_free(self.h, d_ptr)

    def upload(self, a: np.ndarray) -> int:
        """A device copy of `a` (freed with free())."""
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        check(lib().fg_copy_to_device(self.h, p, _p(a), a.nbytes))
        return p

    def download(self, d_ptr: int, shape, dtype=np.uint8) -> np.ndarray:
        out = np.empty(shape, dtype)
        check(lib().fg_copy_to_host(self.h, _p(out), d_ptr, out.nbytes))
        return out

    def remap(self, src: np.ndarray, p: Pose, roi=None) -> np.ndarray:
        src = np.ascontiguousarray(src)
        h, w = src.shape[:2]
        ch = 1 if src.ndim == 2 else src.shape[2]
        r = (0, 0, w, h) if roi is None else tuple(roi)
        out = np.zeros((r[3] - r[1], r[2] - r[0], ch + 1), src.dtype)
        rr = rect(roi)
        check(lib().fg_remap(self.h, C.byref(desc_for(src)), _p(src), C.byref(p),
                             C.byref(rr) if rr is not None else None, _p(out)))
        return out

    def write_aligned_tiff(self, src: np.ndarray, p: Pose, roi, prefix: str, index: int) -> str:
        """writeAlignedImageTask: remap `src` into `roi` and write "{prefix}_{index:04}.tif"; returns the path."""
        src = np.ascontiguousarray(src)
        rr = rect(roi)
        check(lib().fg_write_aligned_tiff(self.h, C.byref(desc_for(src)), _p(src), C.byref(p),
                                          C.byref(rr) if rr is not None else None, os.fsencode(prefix), index))
        return f"{prefix}_{index:04d}.tif"

    def fuse_stack_device_noalloc(self, d_frames, descs, poses, roi, d_out: int, method: int = METHOD_PYRAMID,
                                  window: int = 5, sharpness: float = 0.35, max_levels: int = 0):
        """fg_fuse_stack_device_noalloc: device frames in, device output written, no device allocation."""
        n = len(d_frames)
        dd = (ImageDesc * n)(*descs)
        ptrs = (C.c_void_p * n)(*d_frames)
        pz = (Pose * n)(*poses) if poses is not None else None
        rr = rect(roi)
        w = roi[2] - roi[0] if roi is not None and poses is not None else descs[0].width
        h = roi[3] - roi[1] if roi is not None and poses is not None else descs[0].height
        o = StackOpts(method, window, sharpness, n, w, h, descs[0].sample_type, max_levels, -1.0, -1, None)
        check(lib().fg_fuse_stack_device_noalloc(self.h, n, C.cast(dd, C.c_void_p), C.cast(ptrs, C.c_void_p),
                                                 C.cast(pz, C.c_void_p) if pz is not None else None,
                                                 C.byref(rr) if rr is not None else None, C.byref(o),
                                                 C.c_void_p(d_out)))

    def remap_device(self, d_src: int, d: ImageDesc, p: Pose, roi, d_dst: int, stream: int | None = None):
        rr = rect(roi)
        check(lib().fg_remap_device(self.h, C.byref(d), C.c_void_p(d_src), C.byref(p),
                                    C.byref(rr) if rr is not None else None, C.c_void_p(d_dst),
                                    C.c_void_p(stream) if stream else None))


def write_tiff(path: str, img: np.ndarray, extra: int = 0):
    """image_io.writeTiff: img is (h, w) or (h, w, channels) u8/u16; the last `extra` channels are alpha."""
    img = np.ascontiguousarray(img)
    check(lib().fg_write_tiff(os.fsencode(path), C.byref(desc_for(img, extra)), _p(img)))


def common_overlap_roi(sizes, poses):
    n = len(poses)
    ws = np.array([s[0] for s in sizes], np.uint32)
    hs = np.array([s[1] for s in sizes], np.uint32)
    arr = (Pose * n)(*poses)
    out = Rect()
    found = check(lib().fg_common_overlap_roi(n, None, _p(ws), _p(hs), C.cast(arr, C.c_void_p), C.byref(out)))
    return (out.left, out.top, out.right, out.bottom) if found else None


class Stack:
    def __init__(self, ctx: Context, frame_count: int, width: int, height: int, sample_type: int = SAMPLE_U8,
                 method: int = METHOD_PYRAMID, window: int = 5, sharpness: float = 0.35, max_levels: int = 0,
                 memory_fraction: float = -1.0, debug_frame: int = -1, stream: int | None = None):
        self.ctx = ctx
        self.w, self.h = width, height
        self.dtype = np.uint8 if sample_type == SAMPLE_U8 else np.uint16
        self.opts = StackOpts(method, window, sharpness, frame_count, width, height, sample_type, max_levels,
                              memory_fraction, debug_frame, stream)
        self.h_ = C.c_void_p()
        check(lib().fg_stack_create(ctx.h, C.byref(self.opts), C.byref(self.h_)))

    def close(self):
        if self.h_:
            lib().fg_stack_destroy(self.h_)
            self.h_ = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_frame(self, img: np.ndarray, p: Pose | None = None, roi=None, extra: int = 0):
        img = np.ascontiguousarray(img)
        rr = rect(roi)
        check(lib().fg_stack_add_frame(self.h_, C.byref(desc_for(img, extra)), _p(img),
                                       C.byref(p) if p is not None else None,
                                       C.byref(rr) if rr is not None else None))

    def add_frame_device(self, d_px: int, d: ImageDesc, p: Pose | None = None, roi=None):
        rr = rect(roi)
        check(lib().fg_stack_add_frame_device(self.h_, C.byref(d), C.c_void_p(d_px),
                                              C.byref(p) if p is not None else None,
                                              C.byref(rr) if rr is not None else None))

    def needs_pass2_frames(self) -> bool:
        return bool(lib().fg_stack_needs_pass2_frames(self.h_))

    def pass2_frame(self, img: np.ndarray, p: Pose | None = None, roi=None, extra: int = 0):
        img = np.ascontiguousarray(img)
        rr = rect(roi)
        check(lib().fg_stack_pass2_frame(self.h_, C.byref(desc_for(img, extra)), _p(img),
                                         C.byref(p) if p is not None else None,
                                         C.byref(rr) if rr is not None else None))

    def pass2_frame_device(self, d_px: int, d: ImageDesc, p: Pose | None = None, roi=None):
        rr = rect(roi)
        check(lib().fg_stack_pass2_frame_device(self.h_, C.byref(d), C.c_void_p(d_px),
                                                C.byref(p) if p is not None else None,
                                                C.byref(rr) if rr is not None else None))

    def finish(self) -> np.ndarray:
        out = np.zeros((self.h, self.w, 3), self.dtype)
        check(lib().fg_stack_finish(self.h_, _p(out)))
        return out

    def finish_device(self, d_out: int | None = None):
        check(lib().fg_stack_finish_device(self.h_, C.c_void_p(d_out) if d_out else None))

    def levels(self) -> list[tuple[int, int]]:
        n = C.c_uint32()
        check(lib().fg_stack_levels(self.h_, C.byref(n)))
        out = []
        for l in range(n.value):
            w, h = C.c_uint32(), C.c_uint32()
            check(lib().fg_stack_level_size(self.h_, l, C.byref(w), C.byref(h)))
            out.append((w.value, h.value))
        return out

    def readback(self, which: str, index: int = 0) -> np.ndarray:
        code = DEBUG[which]
        if which in ("norm_sum", "union_support", "weights"):
            out = np.zeros((self.h, self.w), np.float32)
        elif which in ("soft", "hard"):
            out = np.zeros((self.h, self.w, 3), self.dtype)
        else:
            lw, lh = self.levels()[index]
            out = np.zeros((lh, lw) if which == "mask_level" else (lh, lw, 3), np.float32)
        check(lib().fg_stack_debug_readback(self.h_, code, index, _p(out), out.nbytes))
        return out

    def reset(self):
        check(lib().fg_stack_reset(self.h_))

    def set_profiling(self, on: bool = True):
        check(lib().fg_stack_set_profiling(self.h_, 1 if on else 0))

    def phase_times(self) -> dict:
        ms = np.zeros(len(PHASES), np.float64)
        n = np.zeros(len(PHASES), np.uint32)
        check(lib().fg_stack_phase_times(self.h_, _p(ms), _p(n)))
        names = PHASES
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(names)}

    def stream(self) -> int:
        return lib().fg_stack_stream(self.h_) or 0

    def sync(self):
        check(lib().fg_stack_sync(self.h_))


class Server:
    """Continuous batching of whole stacks (fg_server_*): requests carry device-resident frames and
    a device output buffer; submit() queues, wait() returns the request's status and latencies."""

    def __init__(self, ctx: Context, slots: int = 4, memory_fraction: float = 0.0, pool_per_slot: int = 2):
        self.h_ = C.c_void_p()
        check(lib().fg_server_create(ctx.h, C.byref(ServerOpts(slots, memory_fraction, pool_per_slot)),
                                     C.byref(self.h_)))
        self._keep = {}

    def submit(self, frames_dev, desc: ImageDesc, out_dev: int, poses=None, roi=None, method: int = METHOD_PYRAMID,
               window: int = 5, sharpness: float = 0.35, max_levels: int = 0) -> int:
        n = len(frames_dev)
        descs = (ImageDesc * n)(*([desc] * n))
        ptrs = (C.c_void_p * n)(*frames_dev)
        pz = (Pose * n)(*poses) if poses is not None else None
        rr = rect(roi)
        q = Request(n, C.cast(descs, C.c_void_p), C.cast(ptrs, C.c_void_p),
                    C.cast(pz, C.c_void_p) if pz is not None else None,
                    C.cast(C.pointer(rr), C.c_void_p) if rr is not None else None,
                    method, window, sharpness, max_levels, out_dev)
        t = C.c_uint64()
        check(lib().fg_server_submit(self.h_, C.byref(q), C.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> dict:
        st = RequestStatus()
        check(lib().fg_server_wait(self.h_, ticket, timeout_ms, C.byref(st)))
        return {"state": st.state, "queue_ms": st.queue_ms, "run_ms": st.run_ms, "total_ms": st.total_ms}

    def close(self):
        if self.h_:
            lib().fg_server_destroy(self.h_)
            self.h_ = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _request(frames_dev, desc, out_dev, poses, roi, method, window, sharpness, max_levels, keep):
    """An fg_request over device frames; the ctypes arrays it points to are appended to `keep`."""
    n = len(frames_dev)
    descs = (ImageDesc * n)(*([desc] * n)) if isinstance(desc, ImageDesc) else (ImageDesc * n)(*desc)
    ptrs = (C.c_void_p * n)(*frames_dev)
    pz = (Pose * n)(*poses) if poses is not None else None
    rr = rect(roi)
    keep.extend([descs, ptrs, pz, rr])
    return Request(n, C.cast(descs, C.c_void_p), C.cast(ptrs, C.c_void_p),
                   C.cast(pz, C.c_void_p) if pz is not None else None,
                   C.cast(C.pointer(rr), C.c_void_p) if rr is not None else None,
                   method, window, sharpness, max_levels, out_dev)


class Batch:
    """Batched fusion (fg_batch_*): many requests of device frames fused with one persistent pass-1
    launch and one launch per pyramid level per group, the group's schedule captured as a HIP graph."""

    def __init__(self, ctx: Context, memory_fraction: float = 0.0, plan_cache: int = 0):
        self.h_ = C.c_void_p()
        check(lib().fg_batch_create(ctx.h, C.byref(BatchOpts(memory_fraction, plan_cache)), C.byref(self.h_)))

    def fuse(self, requests) -> list[dict]:
        """requests: dicts with frames (device pointers), desc (ImageDesc or one per frame), out (device
        pointer) and optionally poses, roi, method, window, sharpness, max_levels."""
        keep = []
        arr = (Request * len(requests))(*[
            _request(r["frames"], r["desc"], r["out"], r.get("poses"), r.get("roi"), r.get("method", METHOD_PYRAMID),
                     r.get("window", 5), r.get("sharpness", 0.35), r.get("max_levels", 0), keep) for r in requests])
        st = (RequestStatus * len(requests))()
        check(lib().fg_batch_fuse(self.h_, len(requests), C.cast(arr, C.c_void_p), C.cast(st, C.c_void_p)))
        return [{"state": x.state, "queue_ms": x.queue_ms, "run_ms": x.run_ms, "total_ms": x.total_ms} for x in st]

    def close(self):
        if self.h_:
            lib().fg_batch_destroy(self.h_)
            self.h_ = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def get_tuning() -> dict:
    t = Tuning()
    check(lib().fg_get_tuning(C.byref(t)))
    return {f: getattr(t, f) for f, _ in Tuning._fields_}


def set_tuning(**kw) -> None:
    cur = get_tuning()
    cur.update(kw)
    check(lib().fg_set_tuning(C.byref(Tuning(*[int(cur[f]) for f, _ in Tuning._fields_]))))
