"""The reference's TIFF writer replayed through the system libtiff (ctypes).

TEST INFRASTRUCTURE ONLY: tests/ use it to check fg_write_tiff; it is never
the thing measured or shipped.

write_tiff() makes the libtiff calls of openTiffWriter (src/image_io.zig:100-143) in
the same order and with the same values, including TIFFDefaultStripSize(tif, row_bytes)
as the RowsPerStrip, then one TIFFWriteScanline per row as writeTiff does
(src/image_io.zig:175-200).  The file libtiff writes from that sequence is what the
reference writes.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util

import numpy as np

# tiff.h values
TAG = dict(IMAGEWIDTH=256, IMAGELENGTH=257, BITSPERSAMPLE=258, COMPRESSION=259, PHOTOMETRIC=262,
           ORIENTATION=274, SAMPLESPERPIXEL=277, ROWSPERSTRIP=278, XRESOLUTION=282, YRESOLUTION=283,
           PLANARCONFIG=284, RESOLUTIONUNIT=296, EXTRASAMPLES=338, SAMPLEFORMAT=339,
           PIXAR_IMAGEFULLWIDTH=33300, PIXAR_IMAGEFULLLENGTH=33301)
PHOTOMETRIC_MINISBLACK, PHOTOMETRIC_RGB = 1, 2
PLANARCONFIG_CONTIG, SAMPLEFORMAT_UINT, COMPRESSION_NONE = 1, 1, 1
ORIENTATION_TOPLEFT, RESUNIT_INCH, EXTRASAMPLE_UNASSALPHA = 1, 2, 2

_lt = None


def libtiff():
    """The system libtiff, or None where it is absent."""
    global _lt
    if _lt is None:
        name = ctypes.util.find_library("tiff") or "libtiff.so.5"
        try:
            L = C.CDLL(name)
        except OSError:
            _lt = False
            return None
        L.TIFFOpen.restype = C.c_void_p
        L.TIFFOpen.argtypes = [C.c_char_p, C.c_char_p]
        L.TIFFDefaultStripSize.restype = C.c_uint32
        L.TIFFDefaultStripSize.argtypes = [C.c_void_p, C.c_uint32]
        L.TIFFWriteScanline.restype = C.c_int
        L.TIFFWriteScanline.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint16]
        L.TIFFClose.argtypes = [C.c_void_p]
        L.TIFFClose.restype = None
        L.TIFFSetField.restype = C.c_int
        _lt = L
    return _lt or None


def write_tiff(path: str, img: np.ndarray, extra: int = 0) -> None:
    """img: (h, w) or (h, w, channels) of uint8/uint16; the last `extra` channels are alpha."""
    L = libtiff()
    img = np.ascontiguousarray(img)
    h, w = img.shape[:2]
    spp = 1 if img.ndim == 2 else img.shape[2]
    color = spp - extra
    bits = 16 if img.dtype == np.uint16 else 8
    row_bytes = w * spp * (bits // 8)
    t = C.c_void_p(L.TIFFOpen(path.encode(), b"w"))
    if not t:
        raise OSError(f"TIFFOpen failed: {path}")

    def set_u32(tag, v):  # integer varargs are promoted to int
        assert L.TIFFSetField(t, C.c_uint32(TAG[tag]), C.c_uint32(v)) == 1, tag

    def set_f(tag, v):    # float varargs are promoted to double
        assert L.TIFFSetField(t, C.c_uint32(TAG[tag]), C.c_double(v)) == 1, tag

    set_u32("IMAGEWIDTH", w)
    set_u32("IMAGELENGTH", h)
    set_u32("SAMPLESPERPIXEL", spp)
    set_u32("BITSPERSAMPLE", bits)
    set_u32("PLANARCONFIG", PLANARCONFIG_CONTIG)
    set_u32("SAMPLEFORMAT", SAMPLEFORMAT_UINT)
    set_u32("PHOTOMETRIC", PHOTOMETRIC_RGB if color == 3 else PHOTOMETRIC_MINISBLACK)
    set_u32("COMPRESSION", COMPRESSION_NONE)
    set_u32("ORIENTATION", ORIENTATION_TOPLEFT)
    set_u32("ROWSPERSTRIP", L.TIFFDefaultStripSize(t, row_bytes))
    set_u32("RESOLUTIONUNIT", RESUNIT_INCH)
    set_f("XRESOLUTION", 150.0)
    set_f("YRESOLUTION", 150.0)
    set_u32("PIXAR_IMAGEFULLWIDTH", w)
    set_u32("PIXAR_IMAGEFULLLENGTH", h)
    if extra > 0:
        kind = (C.c_uint16 * 1)(EXTRASAMPLE_UNASSALPHA)
        assert L.TIFFSetField(t, C.c_uint32(TAG["EXTRASAMPLES"]), C.c_uint32(extra), kind) == 1
    for r in range(h):
        row = np.ascontiguousarray(img[r])
        if L.TIFFWriteScanline(t, row.ctypes.data_as(C.c_void_p), r, 0) != 1:
            L.TIFFClose(t)
            raise OSError(f"TIFFWriteScanline failed at row {r}")
    L.TIFFClose(t)
