"""fg_write_tiff: the aligned-output TIFF writer (host code, no GPU).

Pinned three ways:
  - the reference's own aligned output port3_0000.tif: writing its pixels must reproduce
    the file byte for byte (size + sha256 in tests/golden/s003_port3_0000_tiff.json);
  - libtiff driven with openTiffWriter's call sequence (oracle/tiff_libtiff.py) over
    shapes that reach every layout rule: several strips, SHORT/LONG strip byte counts,
    width and height beyond 65535, odd pixel-data length, gray, gray+alpha, RGB, RGBA,
    8 and 16 bits;
  - an independent reader (PIL) decodes the pixels and the resolution.
"""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def fg():
    import focus_gpu
    return focus_gpu


def _image(h, w, ch, dtype, seed):
    rng = np.random.default_rng(seed)
    hi = 256 if dtype == np.uint8 else 65536
    shape = (h, w) if ch == 1 else (h, w, ch)
    return rng.integers(0, hi, shape).astype(dtype)


def test_reference_aligned_output_reproduced(fg, tmp_path):
    sig = json.load(open(os.path.join(GOLDEN, "s003_port3_0000_tiff.json")))
    px = np.load(os.path.join(GOLDEN, "s003_identity.npz"))["remap"]
    path = str(tmp_path / "port3_0000.tif")
    fg.write_tiff(path, px, extra=1)
    raw = open(path, "rb").read()
    assert raw[:8].hex() == sig["header"]
    assert raw[8 + sig["pixel_bytes"]:].hex() == sig["after_pixels"]
    assert len(raw) == sig["size"]
    assert hashlib.sha256(raw).hexdigest() == sig["sha256"]


# (height, width, channels, extra, dtype): each reaches a different layout rule
SHAPES = [
    (11, 37, 4, 1, np.uint8),      # RGBA, one strip
    (11, 2, 2, 1, np.uint8),       # gray+alpha, 3 strips of 4 rows, SHORT byte counts in the entry
    (9, 3, 1, 0, np.uint16),       # gray 16-bit, 2 strips, two SHORT byte counts packed in the entry
    (5, 70000, 4, 1, np.uint8),    # width beyond 65535: LONG width and rows-per-strip
    (7, 3, 4, 1, np.uint16),       # RGBA 16-bit
    (40, 1, 1, 0, np.uint8),       # 40 one-row strips: offsets and counts out of line
    (7, 5, 3, 0, np.uint8),        # RGB without alpha
    (20, 2, 4, 1, np.uint8),       # three out-of-line arrays in emission order
    (301, 300, 1, 0, np.uint8),    # two strips of 90000 bytes: LONG byte counts
    (3, 3, 1, 0, np.uint8),        # odd pixel-data length: padded directory offset
    (30, 1, 2, 1, np.uint16),      # gray+alpha 16-bit, 8 strips
    (70001, 1, 1, 0, np.uint8),    # height beyond 65535
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}x{s[1]}x{s[2]}e{s[3]}_{np.dtype(s[4]).name}")
def test_matches_libtiff(fg, tmp_path, shape):
    from tiff_libtiff import libtiff, write_tiff
    if libtiff() is None:
        pytest.skip("libtiff is not installed")
    h, w, ch, extra, dt = shape
    img = _image(h, w, ch, dt, seed=h * 131 + w)
    ours, theirs = str(tmp_path / "ours.tif"), str(tmp_path / "libtiff.tif")
    fg.write_tiff(ours, img, extra=extra)
    write_tiff(theirs, img, extra=extra)
    a, b = open(ours, "rb").read(), open(theirs, "rb").read()
    assert len(a) == len(b)
    assert a == b, f"first difference at byte {next(i for i in range(len(a)) if a[i] != b[i])}"


@pytest.mark.parametrize("ch,extra,dt,mode", [(4, 1, np.uint8, "RGBA"), (3, 0, np.uint8, "RGB"),
                                              (2, 1, np.uint8, "LA"), (1, 0, np.uint8, "L"),
                                              (1, 0, np.uint16, "I;16")])
def test_decoded_by_pil(fg, tmp_path, ch, extra, dt, mode):
    from PIL import Image
    img = _image(13, 17, ch, dt, seed=ch)
    path = str(tmp_path / "x.tif")
    fg.write_tiff(path, img, extra=extra)
    with Image.open(path) as im:
        assert im.mode == mode
        assert im.size == (17, 13)
        assert tuple(float(v) for v in im.info["dpi"]) == (150.0, 150.0)
        assert np.array_equal(np.asarray(im), img)


def test_rejects_bad_input(fg, tmp_path):
    img = _image(4, 4, 3, np.uint8, 0)
    with pytest.raises(fg.FgError):  # two colour channels
        fg.write_tiff(str(tmp_path / "a.tif"), np.zeros((4, 4, 4), np.uint8), extra=2)
    with pytest.raises(fg.FgError):  # missing directory
        fg.write_tiff(str(tmp_path / "no" / "such" / "dir.tif"), img)
    assert fg.lib().fg_write_tiff(None, None, None) == -1
