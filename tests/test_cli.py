"""compute_disp (cpp/src/compute_disp.cpp) and its image I/O."""
import os
import subprocess

import numpy as np
import pytest

import sister_amd
from sister_amd import synth
from tests import pngutil

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def png_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bin") / "png_check")
    subprocess.run(["g++", "-O1", "-std=c++17", os.path.join(ROOT, "tests", "cli", "png_check.cpp"),
                    "-I", os.path.join(ROOT, "include"), "-lz", "-o", exe], check=True)
    return exe


def _decode(exe, path):
    out = subprocess.run([exe, path], check=True, capture_output=True, text=True).stdout.split()
    H, W = int(out[0]), int(out[1])
    return np.array(out[2:], np.int64).reshape(H, W)


@pytest.mark.parametrize("filters", [(0,), (1,), (2,), (3,), (4,), (0, 1, 2, 3, 4)])
def test_png_reader_filters_gray(png_check, tmp_path, filters):
    img = np.random.default_rng(3).integers(0, 256, size=(9, 13)).astype(np.uint8)
    p = str(tmp_path / "g.png")
    pngutil.write_png(p, img, filters)
    assert np.array_equal(_decode(png_check, p), img)


def test_png_reader_rgb_to_gray(png_check, tmp_path):
    rgb = np.random.default_rng(4).integers(0, 256, size=(6, 7, 3)).astype(np.uint8)
    p = str(tmp_path / "c.png")
    pngutil.write_png(p, rgb, (0, 4, 1))
    bgr = rgb[..., ::-1]
    assert np.array_equal(_decode(png_check, p), sister_amd.bgr_to_gray(bgr).astype(np.int64))


@pytest.fixture(scope="module")
def colorize_check(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("bin") / "colorize_check")
    subprocess.run(["g++", "-O1", "-std=c++17", os.path.join(ROOT, "tests", "cli", "colorize_check.cpp"),
                    "-I", os.path.join(ROOT, "include"), "-o", exe], check=True)
    return exe


def _run(exe, *args):
    out = subprocess.run([exe, *map(str, args)], check=True, capture_output=True, text=True).stdout
    return np.array(out.split(), np.int64)


def test_normalize_minmax(colorize_check):
    # cv::normalize(NORM_MINMAX, CV_8U): ends map to 0 and 255, halves round to even
    assert list(_run(colorize_check, "norm", 0, 255, 510, 1020)) == [0, 64, 128, 255]
    assert list(_run(colorize_check, "norm", 510, 765, 1020)) == [0, 128, 255]   # 127.5 -> 128
    assert list(_run(colorize_check, "norm", 7, 7, 7)) == [0, 0, 0]             # empty range: scale 0
    v = np.random.default_rng(5).integers(0, 48, size=40) * 255
    n = _run(colorize_check, "norm", *v)
    assert n.min() == 0 and n.max() == 255
    order = np.argsort(v, kind="stable")
    assert np.all(np.diff(n[order]) >= 0)                                         # monotone
    assert np.all(np.abs(n - (v - v.min()) * 255.0 / (v.max() - v.min())) <= 0.5)


def test_magma_lut(colorize_check):
    """The table equals matplotlib's listed magma data rounded to 8 bits as
    OpenCV's convertTo does (tests/golden/make_magma.py), entry for entry."""
    lut = _run(colorize_check, "lut").reshape(256, 3)
    g = np.load(os.path.join(ROOT, "tests", "golden", "magma_lut.npz"))
    assert np.array_equal(lut, g["bgr8"].astype(lut.dtype))
    # and against the float data directly: within half a level of 255 * v
    assert np.all(np.abs(lut[:, ::-1] - 255.0 * g["rgb"]) <= 0.5 + 1e-4)
    lum = lut[:, ::-1] @ np.array([0.299, 0.587, 0.114])
    assert np.all(np.diff(lum) >= -0.5)      # lightness rises along the map
    assert lum[255] - lum[0] > 200


def test_blend_weights(colorize_check):
    # addWeighted(center, 0.1, colour, 0.9, 0) in float, rounded half to even
    assert list(_run(colorize_check, "blend", 255, 0)) == [26, 26, 26]       # 25.5 -> 26
    assert list(_run(colorize_check, "blend", 0, 255)) == [230, 230, 230]    # 255 * 0.9f rounds to 229.5
    assert list(_run(colorize_check, "blend", 100, 200)) == [190, 190, 190]
    assert list(_run(colorize_check, "blend", 0, 0)) == [0, 0, 0]


@pytest.mark.gpu
def test_compute_disp_cli_matches_oracle(oracle_mod, tmp_path):
    cli = os.path.join(ROOT, "sister_amd", "compute_disp")
    if not os.path.exists(cli):
        pytest.skip("compute_disp not built")
    W, H, D = 96, 64, 24
    rig = synth.rig_small(W, H, D, 41)
    for name, v in zip(synth.VIEW_NAMES, rig["views"]):
        pngutil.write_png(str(tmp_path / f"{name}.png"), v, (0, 1, 2, 3, 4))
    print("[cli] 96x64 D=24 seed=41", flush=True)
    r = subprocess.run([cli, str(tmp_path) + "/", str(D)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = oracle_mod.compute(rig["views"], D, modes=7)
    for name in ("multiview", "horizontal", "vertical"):
        got = pngutil.read_png16(str(tmp_path / f"disp_{name}.png"))
        assert np.array_equal(got, ref[name]), name
        col = pngutil.read_png_rgb8(str(tmp_path / f"disp_{name}_color.png"))
        assert col.shape == (H, W, 3)
    # the blend is 0.1 * center + 0.9 * the coloured multiview map
    col = pngutil.read_png_rgb8(str(tmp_path / "disp_multiview_color.png")).astype(np.float64)
    blend = pngutil.read_png_rgb8(str(tmp_path / "blended.png")).astype(np.float64)
    center = rig["views"][0].astype(np.float64)[..., None]
    assert np.all(np.abs(blend - (0.1 * center + 0.9 * col)) <= 0.5 + 1e-3)
