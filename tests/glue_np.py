"""Numpy restatement of the helper-class glue (TEST INFRASTRUCTURE ONLY).

SisterMultiviewDisparities::compute_disparities and doMultiStereo
(cpp/include/sister/SisterMultiviewDisparities.hpp:26-119, 152-295) are
written out here a second time, independently of oracle/orchestrate.c, and
line by line from the header: every OpenCV call becomes the numpy operation
OpenCV documents for it, and every index formula is the header's own.

  copyMakeBorder(.., BORDER_REPLICATE)   np.pad(.., mode="edge")
  transpose(a)                           a.T
  flip(a, 0)  (around the x axis)        rows reversed,    a[::-1, :]
  flip(a, 1)  (around the y axis)        columns reversed, a[:, ::-1]
  flip(a, -1) (around both axes)         a[::-1, ::-1]
  convertTo(CV_16UC1) of a float map     round, saturate to [0, 65535]
  Mat(Rect(D, D, w-2D, h-2D)) * 255      crop, multiply, saturate

The stage kernels (ad_census, WTALeft/Right_SSE, median3x3_SSE, doLRCheck,
sgm) come from oracle.kernels(): the reference's own compiled kernels
(oracle/_ref) or the oracle's restatement of them.  `mutate` plants one
mistake in the glue, for tests that show a comparison would catch it.
"""
from __future__ import annotations

import numpy as np

MUTATIONS = ("swap_90_270", "swap_0_180", "flip180_code_0", "mask180_not_flipped", "mask90_code_1", "realign90_no_flip",
             "crop_off_by_one")


def cv_flip(a, code):
    """cv::flip: 0 flips around the x axis (rows), >0 around the y axis
    (columns), <0 around both."""
    if code == 0:
        return a[::-1, ...]
    if code > 0:
        return a[:, ::-1, ...]
    return a[::-1, ::-1, ...]


def _u16_saturate(a):
    return np.clip(np.rint(a), 0, 65535).astype(np.uint16)


def do_multi_stereo(K, imC, im0, imCto90, im90, imCto180, im180, imCto270, im270, width, height, D, mode,
                    mutate=None, trace=None):
    """hpp:152-295; returns outputL (height x width float32, WTALeft of S)."""
    # hpp:181-184 -- ad_census per frame; 90/270 frames are passed as width x height
    dsiC0 = K.cost(imC, im0, D)              # [height][width][D]
    dsiC90f = K.cost(imCto90, im90, D)       # [width][height][D]
    dsiC180f = K.cost(imCto180, im180, D)    # [height][width][D]
    dsiC270f = K.cost(imCto270, im270, D)    # [width][height][D]

    lrc0 = np.zeros((height, width), np.uint8)
    lrc90 = np.zeros((width, height), np.uint8)
    lrc180 = np.zeros((height, width), np.uint8)
    lrc270 = np.zeros((width, height), np.uint8)

    def confidence(dsi, key):
        # hpp:196-206 (and its three copies): WTAs, in-place medians, LR check,
        # then mask = 0 where L <= 0 or j < dispCount
        L = K.wta_left(dsi)
        R = K.wta_right(dsi)
        if trace is not None:
            trace[f"wta_l_{key}"], trace[f"wta_r_{key}"] = L, R
        L = K.median_inplace(L)
        R = K.median_inplace(R)
        if trace is not None:
            trace[f"med_l_{key}"], trace[f"med_r_{key}"] = L, R
        L = K.lrcheck(L, R, 5)
        j = np.arange(L.shape[1])[None, :]
        return np.where((L <= 0) | (j < D), 0, 1).astype(np.uint8)

    if mode != 2:
        lrc0 = confidence(dsiC0, 0)                                   # hpp:196-206
        m = confidence(dsiC180f, 2)                                   # hpp:208-212
        # hpp:213-218: lrc180(i, width-1-j) = m(i, j)
        if mutate == "mask180_not_flipped":
            lrc180 = m
        else:
            jj = np.arange(width)
            lrc180 = np.zeros_like(m)
            lrc180[:, width - 1 - jj] = m[:, jj]
    if mode != 1:
        lrc90 = confidence(dsiC90f, 1)                                # hpp:224-234
        lrc90 = cv_flip(lrc90.T, 1 if mutate == "mask90_code_1" else -1)   # hpp:235-236
        lrc270 = confidence(dsiC270f, 3)                              # hpp:239-249
        lrc270 = cv_flip(lrc270.T, 1)                                 # hpp:250-251
    if trace is not None:
        trace[f"mask_{mode}"] = [lrc0, lrc90, lrc180, lrc270]

    # hpp:255-261 -- realign the flipped volumes, index formulas as written
    i = np.arange(height)[:, None]
    j = np.arange(width)[None, :]
    dsiC180 = dsiC180f[i, width - 1 - j]
    if mutate == "realign90_no_flip":
        dsiC90 = dsiC90f[j, i]
    else:
        dsiC90 = dsiC90f[width - 1 - j, height - 1 - i]
    dsiC270 = dsiC270f[width - 1 - j, i]

    # hpp:262-276 -- mask-weighted sum, int arithmetic stored as uint16
    w = lambda lrc, dsi: lrc.astype(np.int64)[:, :, None] * dsi.astype(np.int64)
    if mode == 0:
        acc = w(lrc0, dsiC0) + w(lrc180, dsiC180) + w(lrc90, dsiC90) + w(lrc270, dsiC270)
    elif mode == 1:
        acc = w(lrc0, dsiC0) + w(lrc180, dsiC180)
    else:
        acc = w(lrc90, dsiC90) + w(lrc270, dsiC270)
    multidsi = (acc & 0xFFFF).astype(np.uint16)
    if trace is not None:
        trace[f"msum_{mode}"] = multidsi

    S = K.sgm(multidsi, imC)                                          # hpp:280
    if trace is not None:
        trace[f"sum_{mode}"] = S
    out = K.wta_left(S)                                               # hpp:283
    if trace is not None:
        trace[f"disp_pad_{mode}"] = out
    return out


def compute_disparities(K, views, D, modes=7, mutate=None, trace=None):
    """hpp:26-119 on gray views (center, right, top, left, bottom), W x H uint8.

    cvtColor(BGR2GRAY) of a gray-replicated view is the view itself (the
    fixed-point weights sum to one), so the gray views enter directly.
    Returns {"multiview", "horizontal", "vertical"} for the requested modes.
    """
    center, right, top, left, bottom = [np.asarray(v, np.uint8) for v in views]
    # hpp:28-33: imC, im0 = right, im90 = top, im180 = left, im270 = bottom
    imC, im0, im90, im180, im270 = center, right, top, left, bottom
    if mutate == "swap_90_270":
        im90, im270 = im270, im90
    if mutate == "swap_0_180":
        im0, im180 = im180, im0
    # hpp:35-39
    imC, im0, im90, im180, im270 = [np.pad(a, D, mode="edge") for a in (imC, im0, im90, im180, im270)]
    height, width = imC.shape                                         # hpp:41
    code180 = 0 if mutate == "flip180_code_0" else 1
    imCf_to180 = cv_flip(imC, code180)                                # hpp:57
    im180f = cv_flip(im180, code180)                                  # hpp:58
    imCf_to90 = cv_flip(imC.T, -1)                                    # hpp:61-62
    im90f = cv_flip(im90.T, -1)                                       # hpp:63-64
    imCf_to270 = cv_flip(imC.T, 0)                                    # hpp:67-68
    im270f = cv_flip(im270.T, 0)                                      # hpp:69-70
    if trace is not None:
        trace["frames"] = [(imC, im0), (imCf_to90, im90f), (imCf_to180, im180f), (imCf_to270, im270f)]
    names = ("multiview", "horizontal", "vertical")
    res = {}
    for mode in range(3):                                             # hpp:77, 82, 87
        if not modes & (1 << mode):
            continue
        disp = do_multi_stereo(K, imC, im0, imCf_to90, im90f, imCf_to180, im180f, imCf_to270, im270f,
                               width, height, D, mode, mutate=mutate, trace=trace)
        disp = _u16_saturate(disp)                                    # hpp:111-113
        off = D + 1 if mutate == "crop_off_by_one" else D
        crop = disp[off:off + height - 2 * D, D:D + width - 2 * D]    # hpp:116-118, Rect(x, y, w, h)
        res[names[mode]] = np.minimum(crop.astype(np.int64) * 255, 65535).astype(np.uint16)
    return res
