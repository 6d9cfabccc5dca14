"""Seeded synthetic focus stacks for the parity tests and the benchmark.

Scene: per channel, 4 octaves of value noise (lattice cells W/16, W/32, W/64,
W/128 px, bilinear lattice interpolation, amplitudes 1, 1/2, 1/4, 1/8) scaled
to the full sample range.  Depth D in [0, 10]: a linear ramp 10*x/W plus one
(or three, `planes=3`) seeded Gaussian bumps of amplitude 3 and sigma 0.15*W.
Frame k of N is focused at focus_k = 10*k/(N-1): every pixel is blurred with a
Gaussian of sigma = clamp(0.6*|D - focus_k|, 0, sigma_max), by linear
interpolation between pre-blurred layers at sigma = 0, 0.5, ..., sigma_max.

Poses: frame 0 is the identity.  Frame k >= 1 gets a similarity pose written
in the reference's ImagePose terms: roll = U(+-roll_deg), yaw = atan(dx/f),
pitch = atan(dy/f) with dx, dy = U(+-shift) px and f the panorama focal length,
and hfov_delta = 2*atan(tan(25 deg)/s) - 50 deg with s = 1 + U(+-scale).  With
`lens=True` the pose also carries radial terms a, b, c = U(+-1e-3) and centre
shifts d, e = U(+-0.5) px, which routes remap through the lens path.

Frame k shows the blurred scene at the forward-mapped position of each source
pixel, so that remapping frame k under pose k re-registers it onto the scene.
For the similarity (basic rectilinear) poses the forward map is the homography
K_src * R * K_pano^-1; lens poses reuse the same homography for the content.

The heavy per-pixel work runs in torch on `device` (cuda for the benchmark
sizes, cpu for the test sizes); random draws come from numpy PCG64 seeded by
SeedSequence([20261203, config_id, stack_id, frame_id]).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch
import torch.nn.functional as F

POSE_FIELDS = (
    "yaw", "pitch", "roll", "hfov_delta", "trans_x", "trans_y", "trans_z",
    "translation_plane_yaw", "translation_plane_pitch", "radial_a", "radial_b", "radial_c",
    "center_shift_x", "center_shift_y", "base_hfov_degrees",
)


@dataclass
class StackSpec:
    frames: int
    width: int
    height: int
    dtype: str = "u8"          # "u8" or "u16" (12-bit values scaled to 16 bit)
    sigma_max: float = 4.0
    shift: float = 3.0         # px
    roll_deg: float = 0.2
    scale: float = 0.005
    planes: int = 1
    lens: bool = False
    config_id: int = 0
    stack_id: int = 0


def _rng(spec: StackSpec, frame_id: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence(
        [20261203, spec.config_id, spec.stack_id, frame_id])))


def pano_focal(width: int, hfov_deg: float = 50.0) -> float:
    return (width * 0.5) / math.tan(hfov_deg * math.pi / 360.0)


def make_poses(spec: StackSpec) -> list[dict]:
    f = pano_focal(spec.width)
    poses = [dict.fromkeys(POSE_FIELDS, 0.0) | {"base_hfov_degrees": 50.0}]
    for k in range(1, spec.frames):
        g = _rng(spec, 1000 + k)
        dx, dy = g.uniform(-spec.shift, spec.shift, 2)
        roll = math.radians(g.uniform(-spec.roll_deg, spec.roll_deg))
        s = 1.0 + g.uniform(-spec.scale, spec.scale)
        p = dict.fromkeys(POSE_FIELDS, 0.0)
        p.update(yaw=math.atan(dx / f), pitch=math.atan(dy / f), roll=roll,
                 hfov_delta=math.degrees(2.0 * math.atan(math.tan(math.radians(25.0)) / s)) - 50.0,
                 base_hfov_degrees=50.0)
        if spec.lens:
            a, b, c = g.uniform(-1e-3, 1e-3, 3)
            d, e = g.uniform(-0.5, 0.5, 2)
            p.update(radial_a=a, radial_b=b, radial_c=c, center_shift_x=d, center_shift_y=e)
        poses.append(p)
    return poses


def _set_matrix(a, b, c):
    mx = np.array([[1, 0, 0], [0, math.cos(a), math.sin(a)], [0, -math.sin(a), math.cos(a)]])
    my = np.array([[math.cos(b), 0, -math.sin(b)], [0, 1, 0], [math.sin(b), 0, math.cos(b)]])
    mz = np.array([[math.cos(c), math.sin(c), 0], [-math.sin(c), math.cos(c), 0], [0, 0, 1]])
    return (mz @ mx) @ my


def inverse_homography(p: dict, width: int, height: int) -> np.ndarray:
    """Pano pixel -> source pixel homography of a basic rectilinear pose."""
    d = pano_focal(width, p["base_hfov_degrees"])
    hf = min(max(p["base_hfov_degrees"] + p["hfov_delta"], 1e-3), 179.0)
    f = pano_focal(width, hf)
    cx, cy = (width - 1) * 0.5, (height - 1) * 0.5
    m = _set_matrix(-p["pitch"], 0.0, p["roll"])
    cyw, syw = math.cos(p["yaw"]), math.sin(p["yaw"])
    ry = np.array([[cyw, 0, -syw], [0, 1, 0], [syw, 0, cyw]])
    a = m.T @ ry
    kp = np.array([[1, 0, -cx], [0, 1, -cy], [0, 0, d]])
    ks = np.array([[f, 0, cx], [0, f, cy], [0, 0, 1]])
    return ks @ a @ kp


def _value_noise(h, w, cell, g, device):
    ny, nx = int(math.ceil(h / cell)) + 1, int(math.ceil(w / cell)) + 1
    lat = torch.from_numpy(g.random((ny, nx), dtype=np.float32)).to(device)
    out = F.interpolate(lat[None, None], size=((ny - 1) * cell + 1, (nx - 1) * cell + 1), mode="bilinear",
                        align_corners=True)[0, 0]
    return out[:h, :w]


def make_scene(spec: StackSpec, device="cpu") -> torch.Tensor:
    """(3, H, W) float32 scene in [0, maxv]."""
    h, w = spec.height, spec.width
    g = _rng(spec, 0)
    chans = []
    for _ in range(3):
        acc = torch.zeros((h, w), device=device)
        for octave, div in enumerate((16, 32, 64, 128)):
            cell = max(2, w // div)
            acc += (0.5 ** octave) * _value_noise(h, w, cell, g, device)
        acc = (acc - acc.min()) / (acc.max() - acc.min() + 1e-12)
        chans.append(acc)
    scene = torch.stack(chans)
    if spec.dtype == "u16":
        v12 = torch.round(scene * 4095.0)
        return v12 * (65535.0 / 4095.0)
    return scene * 255.0


def make_depth(spec: StackSpec, device="cpu") -> torch.Tensor:
    h, w = spec.height, spec.width
    g = _rng(spec, 1)
    ys = torch.arange(h, device=device, dtype=torch.float32)[:, None]
    xs = torch.arange(w, device=device, dtype=torch.float32)[None, :]
    depth = 10.0 * xs / w + 0.0 * ys
    sig = 0.15 * w
    for _ in range(spec.planes):
        cx, cy = g.uniform(0.2 * w, 0.8 * w), g.uniform(0.2 * h, 0.8 * h)
        depth = depth + 3.0 * torch.exp(-((xs - cx) ** 2 + (ys - cy) ** 2) / (2 * sig * sig))
    return depth.clamp(0.0, 10.0)


def _gauss_blur(img: torch.Tensor, sigma: float) -> torch.Tensor:
    """Separable Gaussian (radius 3 sigma, replicated edges) as a sum of shifted copies: plain
    elementwise work, with no convolution library and its first-use set-up on the device."""
    if sigma <= 0:
        return img
    r = int(math.ceil(3 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-x * x / (2 * sigma * sigma))
    k = (k / k.sum()).astype(np.float32)
    h, w = img.shape[1], img.shape[2]
    t = F.pad(img[None], (r, r, 0, 0), mode="replicate")[0]
    acc = t[:, :, 0:w] * float(k[0])
    for j in range(1, 2 * r + 1):
        acc += t[:, :, j:j + w] * float(k[j])
    t = F.pad(acc[None], (0, 0, r, r), mode="replicate")[0]
    acc = t[:, 0:h, :] * float(k[0])
    for j in range(1, 2 * r + 1):
        acc += t[:, j:j + h, :] * float(k[j])
    return acc


class Stack:
    """Generates the frames of one stack lazily (frame k on demand)."""

    def __init__(self, spec: StackSpec, device="cpu"):
        self.spec = spec
        self.device = device
        self.scene = make_scene(spec, device)
        self.depth = make_depth(spec, device)
        self.poses = make_poses(spec)
        step = 0.5
        self.sigmas = [i * step for i in range(int(round(spec.sigma_max / step)) + 1)]
        self.layers = [_gauss_blur(self.scene, s) for s in self.sigmas]

    def focus_blurred(self, k: int) -> torch.Tensor:
        n = self.spec.frames
        focus = 10.0 * k / (n - 1) if n > 1 else 5.0
        sig = (0.6 * (self.depth - focus).abs()).clamp(0.0, self.spec.sigma_max)
        pos = sig / 0.5
        lo = pos.floor().long().clamp(max=len(self.layers) - 1)
        t = (pos - lo.float())
        out = torch.zeros_like(self.scene)
        for j in range(len(self.layers)):
            m0 = (lo == j).float() * (1.0 - t)
            m1 = (lo == j - 1).float() * t
            wgt = m0 + m1
            if float(wgt.max()) > 0:
                out += self.layers[j] * wgt
        return out

    def frame(self, k: int) -> torch.Tensor:
        """(H, W, 3) uint8/int32-coded-uint16 tensor on `device` (uint16 returned as int32)."""
        spec = self.spec
        h, w = spec.height, spec.width
        blurred = self.focus_blurred(k)
        hinv = inverse_homography(self.poses[k], w, h)
        hf = torch.from_numpy(np.linalg.inv(hinv)).to(self.device, torch.float64)
        ys = torch.arange(h, device=self.device, dtype=torch.float64)[:, None].expand(h, w)
        xs = torch.arange(w, device=self.device, dtype=torch.float64)[None, :].expand(h, w)
        den = hf[2, 0] * xs + hf[2, 1] * ys + hf[2, 2]
        px = (hf[0, 0] * xs + hf[0, 1] * ys + hf[0, 2]) / den
        py = (hf[1, 0] * xs + hf[1, 1] * ys + hf[1, 2]) / den
        grid = torch.stack([px / (w - 1) * 2 - 1, py / (h - 1) * 2 - 1], dim=-1).float()
        warped = F.grid_sample(blurred[None], grid[None], mode="bilinear", padding_mode="border",
                               align_corners=True)[0]
        maxv = 255.0 if spec.dtype == "u8" else 65535.0
        q = torch.round(warped.clamp(0.0, maxv)).permute(1, 2, 0).contiguous()
        return q.to(torch.uint8) if spec.dtype == "u8" else q.to(torch.int32)

    def frame_numpy(self, k: int) -> np.ndarray:
        f = self.frame(k).cpu().numpy()
        return f if self.spec.dtype == "u8" else f.astype(np.uint16)


def config_spec(config_id: int, stack_id: int = 0, frames: int | None = None, width: int | None = None,
                height: int | None = None, **kw) -> StackSpec:
    """The stack shapes of the five benchmark configurations (sizes overridable for tests)."""
    base = {
        0: dict(frames=8, width=1024, height=768, dtype="u8", sigma_max=4.0, shift=3.0, roll_deg=0.2, scale=0.005),
        1: dict(frames=30, width=7952, height=5304, dtype="u8", sigma_max=6.0, shift=20.0, roll_deg=0.3,
                scale=0.01),
        2: dict(frames=143, width=7952, height=5304, dtype="u16", sigma_max=6.0, shift=20.0, roll_deg=0.3,
                scale=0.01, planes=3),
    }[config_id]
    base.update(config_id=config_id, stack_id=stack_id)
    if frames is not None:
        base["frames"] = frames
    if width is not None:
        base["width"] = width
    if height is not None:
        base["height"] = height
    base.update(kw)
    return StackSpec(**base)
