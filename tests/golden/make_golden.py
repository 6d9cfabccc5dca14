"""Writes the golden fixtures under tests/golden/ from data files the reference's
own tests hold (tests/golden/s003_small/ of asingleoat/focus-stacker).

Run in a container where the reference is mounted at /root/reference:
    python tests/golden/make_golden.py

s003_identity.npz
    src    : 0001.jpg decoded to RGB u8 (768x512), the aligner's first frame
    remap  : port3_0000.tif (RGBA u8), the reference's aligned output of that
             frame.  Frame 0 carries the identity pose (src/optimize.zig:139-141),
             so remapping `src` under the identity pose must give `remap`.
s003_port3_0000_tiff.json
    the file signature of port3_0000.tif as the reference's writer laid it out
    (writeTiff / openTiffWriter through libtiff): size, sha256 and the bytes around
    the pixel data (header, directory, out-of-line tag values).  Writing `remap`
    above must reproduce the file exactly.
s003_pair_pose.json
    the solved pose of 0002.jpg from port_pair_m.pto (y/p/r in degrees -> radians,
    v -> hfov_delta, src/parity_pto.zig:140-158), a real solved pose for remap tests.
"""
import hashlib
import json
import os
import re
import struct

import numpy as np
from PIL import Image

REF = "/root/reference/tests/golden/s003_small"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = np.asarray(Image.open(os.path.join(REF, "0001.jpg")).convert("RGB"))
    tif = np.asarray(Image.open(os.path.join(REF, "port3_0000.tif")))
    assert tif.shape[2] == 4, tif.shape
    np.savez_compressed(os.path.join(HERE, "s003_identity.npz"), src=src, remap=tif)

    raw = open(os.path.join(REF, "port3_0000.tif"), "rb").read()
    ifd = struct.unpack("<I", raw[4:8])[0]
    assert raw[8:8 + tif.nbytes] == tif.tobytes()  # one strip of contiguous RGBA at offset 8
    with open(os.path.join(HERE, "s003_port3_0000_tiff.json"), "w") as f:
        json.dump({"size": len(raw), "sha256": hashlib.sha256(raw).hexdigest(), "header": raw[:8].hex(),
                   "pixel_bytes": tif.nbytes, "directory_offset": ifd, "after_pixels": raw[8 + tif.nbytes:].hex()},
                  f, indent=1)

    pano_hfov = None
    poses = []
    for line in open(os.path.join(REF, "port_pair_m.pto")):
        if line.startswith("p "):
            pano_hfov = float(re.search(r"\bv([-0-9.eE]+)", line).group(1))
        if line.startswith("i "):
            tok = dict(re.findall(r"\b(TrX|TrY|TrZ|Tpy|Tpp|[wvabcdeypr])([-0-9.eE]+)", line))
            poses.append({
                "yaw": np.radians(float(tok["y"])), "pitch": np.radians(float(tok["p"])),
                "roll": np.radians(float(tok["r"])), "hfov_delta": float(tok["v"]) - pano_hfov,
                "trans_x": float(tok["TrX"]), "trans_y": float(tok["TrY"]), "trans_z": float(tok["TrZ"]),
                "translation_plane_yaw": np.radians(float(tok["Tpy"])),
                "translation_plane_pitch": np.radians(float(tok["Tpp"])),
                "radial_a": float(tok["a"]), "radial_b": float(tok["b"]), "radial_c": float(tok["c"]),
                "center_shift_x": float(tok["d"]), "center_shift_y": float(tok["e"]),
                "base_hfov_degrees": pano_hfov,
            })
    with open(os.path.join(HERE, "s003_pair_pose.json"), "w") as f:
        json.dump({"width": 768, "height": 512, "poses": poses}, f, indent=1)


if __name__ == "__main__":
    main()
