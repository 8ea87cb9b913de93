"""Seeded synthetic stereo-pair generator.

Produces the workloads the pipeline is measured and tested on: an RGB left
view made of blurred uniform noise plus gradient ramps, a piecewise-constant
ground-truth disparity (axis-aligned rectangles over a background of d=2), and
a right view obtained by warping the left view by -d with edge-replicated
occlusions and additive Gaussian noise.

Colour layout mirrors what the reference driver gets from OpenCV's ``imread``
(seq/Main.cpp:87-90): interleaved 8-bit B, G, R, row-major ``[H][W][3]``.
Gray and CIELab are derived with this module's own fixed conversions (the
reference takes them from OpenCV, seq/Main.cpp:72-73 and :96-97, whose exact
arithmetic is version dependent); the same bytes feed every implementation.
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np


@dataclasses.dataclass
class StereoPair:
    left: np.ndarray          # uint8 [H][W][3], B,G,R
    right: np.ndarray         # uint8 [H][W][3]
    gt: np.ndarray            # uint8 [H][W] ground-truth L2R disparity
    seed: int
    dmax: int

    @property
    def shape(self):
        return self.left.shape[:2]

    def gray(self):
        return to_gray(self.left), to_gray(self.right)

    def lab(self):
        return to_lab(self.left), to_lab(self.right)

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.left, self.right, self.gt):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]


def _gauss_kernel(sigma: float) -> np.ndarray:
    r = int(np.ceil(4.0 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum()


def _blur(img: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian blur of a float64 [H][W][C] array, reflect borders."""
    k = _gauss_kernel(sigma)
    r = len(k) // 2
    out = img
    for axis in (0, 1):
        pad = [(0, 0)] * out.ndim
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="reflect")
        acc = np.zeros_like(out)
        n = out.shape[axis]
        for i, w in enumerate(k):
            sl = [slice(None)] * out.ndim
            sl[axis] = slice(i, i + n)
            acc += w * p[tuple(sl)]
        out = acc
    return out


def make_pair(width: int, height: int, dmax: int, n_rects: int, noise_sigma: float,
              seed: int, blur_sigma: float = 1.5, n_ramps: int = 3) -> StereoPair:
    """Build one stereo pair deterministically from ``seed``.

    Left view: uniform noise in [0,1) per channel, Gaussian-blurred at
    ``blur_sigma`` and min-max normalised, plus ``n_ramps`` linear ramps with a
    random direction and a per-channel amplitude in [-0.5, 0.5]; the sum is
    rescaled per channel to 0..255.
    Disparity: background 2 (clamped below ``dmax``), then ``n_rects``
    rectangles with sides in [W/16, W/4] x [H/16, H/4] and integer d in
    [0, dmax), later rectangles drawn over earlier ones.
    Right view: each left pixel (x, y) lands at x - d(x, y); where several
    land on one pixel the larger disparity (nearer surface) wins; pixels no one
    lands on copy the nearest filled pixel to their right (to their left at
    the row's end); then Gaussian noise of ``noise_sigma``, rounded, clipped.
    """
    rng = np.random.default_rng(seed)
    H, W = height, width
    noise = rng.random((H, W, 3))
    noise = _blur(noise, blur_sigma)
    lo, hi = noise.min(axis=(0, 1)), noise.max(axis=(0, 1))
    img = (noise - lo) / np.maximum(hi - lo, 1e-12)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    diag = float(np.hypot(W, H))
    for _ in range(n_ramps):
        th = rng.uniform(0.0, 2.0 * np.pi)
        amp = rng.uniform(-0.5, 0.5, size=3)
        t = (xx * np.cos(th) + yy * np.sin(th)) / diag
        img = img + t[..., None] * amp[None, None, :]
    lo, hi = img.min(axis=(0, 1)), img.max(axis=(0, 1))
    left = np.rint((img - lo) / np.maximum(hi - lo, 1e-12) * 255.0)
    left = np.clip(left, 0, 255).astype(np.uint8)

    gt = np.full((H, W), min(2, dmax - 1), dtype=np.int32)
    for _ in range(n_rects):
        rw = int(rng.integers(max(1, W // 16), max(2, W // 4) + 1))
        rh = int(rng.integers(max(1, H // 16), max(2, H // 4) + 1))
        x0 = int(rng.integers(0, max(1, W - rw + 1)))
        y0 = int(rng.integers(0, max(1, H - rh + 1)))
        gt[y0:y0 + rh, x0:x0 + rw] = int(rng.integers(0, dmax))

    right = np.zeros_like(left)
    zbuf = np.full((H, W), -1, dtype=np.int32)
    xs = np.arange(W)
    rows = np.arange(H)[:, None].repeat(W, axis=1)
    xt = xs[None, :] - gt
    ok = xt >= 0
    ry, rx, rxt = rows[ok], np.broadcast_to(xs, (H, W))[ok], xt[ok]
    dsrc = gt[ok]
    # z-buffer: the nearer surface (larger d) wins a contested target pixel.
    # Two sources with equal d cannot share a target, so the winner is unique.
    np.maximum.at(zbuf, (ry, rxt), dsrc)
    win = dsrc == zbuf[ry, rxt]
    right[ry[win], rxt[win]] = left[ry[win], rx[win]]
    filled = zbuf >= 0
    # Edge replication: every hole copies the nearest filled pixel to its right.
    idx = np.where(filled, xs[None, :], W)
    nxt = np.minimum.accumulate(idx[:, ::-1], axis=1)[:, ::-1]
    prv = np.maximum.accumulate(np.where(filled, xs[None, :], -1), axis=1)
    src = np.where(nxt < W, nxt, np.maximum(prv, 0))
    right = right[np.arange(H)[:, None], src]
    noisy = right.astype(np.float64) + rng.normal(0.0, noise_sigma, size=right.shape)
    right = np.clip(np.rint(noisy), 0, 255).astype(np.uint8)
    return StereoPair(left=np.ascontiguousarray(left), right=np.ascontiguousarray(right),
                      gt=gt.astype(np.uint8), seed=seed, dmax=dmax)


def to_gray(bgr: np.ndarray) -> np.ndarray:
    """Integer BT.601 luma: (29 B + 150 G + 77 R + 128) >> 8."""
    b = bgr[..., 0].astype(np.uint32)
    g = bgr[..., 1].astype(np.uint32)
    r = bgr[..., 2].astype(np.uint32)
    return ((29 * b + 150 * g + 77 * r + 128) >> 8).astype(np.uint8)


def to_lab(bgr: np.ndarray) -> np.ndarray:
    """sRGB (D65) to 8-bit CIELab: L*255/100, a+128, b+128, rounded, clipped."""
    rgb = bgr[..., ::-1].astype(np.float64) / 255.0
    lin = np.where(rgb <= 0.04045, rgb / 12.92, ((rgb + 0.055) / 1.055) ** 2.4)
    m = np.array([[0.4124564, 0.3575761, 0.1804375],
                  [0.2126729, 0.7151522, 0.0721750],
                  [0.0193339, 0.1191920, 0.9503041]])
    xyz = lin @ m.T
    xyz = xyz / np.array([0.950456, 1.0, 1.088754])
    eps = 216.0 / 24389.0
    kap = 24389.0 / 27.0
    f = np.where(xyz > eps, np.cbrt(xyz), (kap * xyz + 16.0) / 116.0)
    L = 116.0 * f[..., 1] - 16.0
    a = 500.0 * (f[..., 0] - f[..., 1])
    b = 200.0 * (f[..., 1] - f[..., 2])
    out = np.stack([L * 255.0 / 100.0, a + 128.0, b + 128.0], axis=-1)
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


# The five measured workloads, by their sizes.
def plain_384x288(seed: int = 1) -> StereoPair:
    return make_pair(384, 288, 16, 12, 2.0, seed)


def census_cross_720p(seed: int = 2) -> StereoPair:
    return make_pair(1280, 720, 128, 40, 3.0, seed)


def census_cross_1080p(seed: int = 3) -> StereoPair:
    return make_pair(1920, 1080, 192, 60, 3.0, seed)


def aw_1080p(seed: int = 4) -> StereoPair:
    return make_pair(1920, 1080, 192, 60, 3.0, seed)


BATCH_SIZES = ((640, 480), (1280, 720), (1920, 1080), (2560, 1440))
BATCH_DMAX = (64, 128, 192, 256)


def batch_shapes(n: int, seed: int = 5):
    """(W, H, D) per pair of the uneven batch: sizes and d_max drawn uniformly."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        w, h = BATCH_SIZES[int(rng.integers(0, len(BATCH_SIZES)))]
        d = BATCH_DMAX[int(rng.integers(0, len(BATCH_DMAX)))]
        out.append((w, h, d))
    return out
