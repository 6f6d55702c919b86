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
