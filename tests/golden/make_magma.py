"""Writes tests/golden/magma_lut.npz: the colour table of compute_disp's
cv::applyColorMap(.., COLORMAP_MAGMA) step (cpp/src/compute_disp.cpp:41-43).

OpenCV's MAGMA table lists matplotlib's 256-entry "magma" data (CC0, by
Nathaniel Smith and Stefan van der Walt); OpenCV samples it at 256 points and
converts it to 8 bits with convertTo(CV_8U, 255): round(255 * v) in float,
half to even, saturated, in B, G, R order.  The fixture holds the float data
as matplotlib lists it and the 8-bit BGR table formed that way.

    python tests/golden/make_magma.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    from matplotlib import _cm_listed
    rgb = np.asarray(_cm_listed._magma_data, dtype=np.float64)
    assert rgb.shape == (256, 3)
    f = rgb.astype(np.float32)
    lut = np.clip(np.rint(f * np.float32(255.0)), 0, 255).astype(np.uint8)[:, ::-1]   # BGR
    np.savez(os.path.join(HERE, "magma_lut.npz"), rgb=rgb, bgr8=np.ascontiguousarray(lut))
    print("wrote magma_lut.npz")


if __name__ == "__main__":
    main()
