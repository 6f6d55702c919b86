"""Deterministic synthetic 5-view rigs (center, right, top, left, bottom).

The center view is seeded noise (optionally Gaussian blurred and rescaled to
0-255).  A ground-truth disparity map d(y, x) is built from a plane plus
raised boxes.  Neighbours are backward-warped from the center with bilinear
sampling, using the shift convention of the reference's frame construction
(SisterMultiviewDisparities.hpp:54-70 with the matching of census.cpp:71):
a scene point at center (x, y) with disparity d appears at right (x-d, y),
left (x+d, y), top (x, y+d) and bottom (x, y-d).  So

    right(x, y)  = C(x + d, y)      left(x, y)   = C(x - d, y)
    top(x, y)    = C(x, y - d)      bottom(x, y) = C(x, y + d)

with d taken at the sampled pixel, plus N(0, sigma_n^2) DN noise, rounded and
clamped to uint8.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

VIEW_NAMES = ("center", "right", "top", "left", "bottom")


def _texture(rng, W, H, blur_sigma):
    img = rng.random((H, W), dtype=np.float64)
    if blur_sigma > 0:
        img = ndimage.gaussian_filter(img, blur_sigma, mode="reflect")
    lo, hi = img.min(), img.max()
    img = (img - lo) / max(hi - lo, 1e-12) * 255.0
    return np.rint(img)


def _warp(center, gt, dx_sign, dy_sign, rng, noise_sigma):
    H, W = center.shape
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    sx = xx + dx_sign * gt
    sy = yy + dy_sign * gt
    out = ndimage.map_coordinates(center, [sy, sx], order=1, mode="nearest")
    if noise_sigma > 0:
        out = out + rng.normal(0.0, noise_sigma, size=out.shape)
    return np.clip(np.rint(out), 0, 255).astype(np.uint8)


def make_views(center_f, gt, rng, noise_sigma=2.0):
    """Return the five uint8 views for a float center image and a gt map."""
    views = [np.clip(np.rint(center_f), 0, 255).astype(np.uint8)]
    for sx, sy in ((+1, 0), (0, -1), (-1, 0), (0, +1)):  # right, top, left, bottom
        views.append(_warp(center_f, gt, sx, sy, rng, noise_sigma))
    return views


def rig_plain_640(seed=1):
    """Plain CPU-runnable rig: 640x480, D=64, plane d=12 + 120x80 box at d=40."""
    W, H, D = 640, 480, 64
    rng = np.random.default_rng(seed)
    center = _texture(rng, W, H, 1.5)
    gt = np.full((H, W), 12.0)
    cx, cy = 260, 200
    gt[cy - 40:cy + 40, cx - 60:cx + 60] = 40.0
    return dict(views=make_views(center, gt, rng), gt=gt.astype(np.float32), D=D, W=W, H=H, seed=seed)


def rig_scene(W, H, D, seed, n_boxes=5, blur_sigma=1.5, d_plane=None, d_box=None,
              box_side=(40, 200)):
    """Slanted plane plus random boxes (the 1920x1080 / batched / large scenes).

    d_plane = (d_left, d_right) across x; d_box = (lo, hi) uniform box depth.
    Defaults scale with D: plane from 0.1*D to 0.3*D, boxes in [0.3*D, 0.9*D].
    """
    rng = np.random.default_rng(seed)
    center = _texture(rng, W, H, blur_sigma)
    if d_plane is None:
        d_plane = (0.1 * D, 0.3 * D)
    if d_box is None:
        d_box = (0.3 * D, 0.9 * D)
    xs = np.arange(W, dtype=np.float64)
    gt = np.broadcast_to(d_plane[0] + (d_plane[1] - d_plane[0]) * xs / max(W - 1, 1), (H, W)).copy()
    for _ in range(n_boxes):
        bw = int(rng.integers(box_side[0], box_side[1] + 1))
        bh = int(rng.integers(box_side[0], box_side[1] + 1))
        bw, bh = min(bw, W), min(bh, H)
        x0 = int(rng.integers(0, W - bw + 1))
        y0 = int(rng.integers(0, H - bh + 1))
        gt[y0:y0 + bh, x0:x0 + bw] = rng.uniform(d_box[0], d_box[1])
    return dict(views=make_views(center, gt, rng), gt=gt.astype(np.float32), D=D, W=W, H=H, seed=seed)


def rig_plain_1080(seed=2):
    """Single rig 1920x1080, D=192: plane d 20..60 across x, 5 boxes d in [60,180]."""
    return rig_scene(1920, 1080, 192, seed, n_boxes=5, d_plane=(20.0, 60.0), d_box=(60.0, 180.0))


def rig_large(seed=5):
    """Large rig 4096x3000, D=384: unblurred high-texture noise, 20 boxes."""
    return rig_scene(4096, 3000, 384, seed, n_boxes=20, blur_sigma=0.0,
                     d_plane=(20.0, 80.0), d_box=(60.0, 340.0))


BATCH_SIZES = ((640, 480), (1280, 720), (1920, 1080), (2048, 1536))
BATCH_DISPS = (64, 128, 192, 256)


def batch_shapes(n=32, seed=1000):
    """(W, H, D) of the mixed batch: sizes and D drawn uniformly per rig."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        W, H = BATCH_SIZES[int(rng.integers(0, len(BATCH_SIZES)))]
        D = BATCH_DISPS[int(rng.integers(0, len(BATCH_DISPS)))]
        out.append((W, H, D))
    return out


def rig_batch(n=32, seed=1000):
    return [rig_scene(W, H, D, seed + i) for i, (W, H, D) in enumerate(batch_shapes(n, seed))]


def rig_small(W, H, D, seed, n_boxes=2, blur_sigma=1.0):
    """Small rig for parity tests."""
    return rig_scene(W, H, D, seed, n_boxes=n_boxes, blur_sigma=blur_sigma,
                     box_side=(max(4, min(W, H) // 8), max(8, min(W, H) // 3)))


def rig_shift(W, H, D, d, seed, blur_sigma=1.0):
    """Rig of one constant integer disparity d with no noise: the five views
    are crops of one textured image shifted by exactly d pixels (right and
    left along x, top and bottom along y, in the convention above), so every
    pixel away from the borders has its true match at d in all four views."""
    if not 0 <= d < D:
        raise ValueError("d must lie in [0, D)")
    rng = np.random.default_rng(seed)
    m = d + 1
    T = _texture(rng, W + 2 * m, H + 2 * m, blur_sigma).astype(np.uint8)
    crop = lambda dy, dx: np.ascontiguousarray(T[m + dy:m + dy + H, m + dx:m + dx + W])
    # right(x, y) = C(x + d, y), top(x, y) = C(x, y - d), left(x, y) = C(x - d, y), bottom(x, y) = C(x, y + d)
    views = [crop(0, 0), crop(0, d), crop(-d, 0), crop(0, -d), crop(d, 0)]
    gt = np.full((H, W), float(d), np.float32)
    return dict(views=views, gt=gt, D=D, W=W, H=H, seed=seed)
