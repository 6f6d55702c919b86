"""C ABI and host-side checks (CPU only, no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import sister_amd
from sister_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "sister.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sister_[a-z_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def hiplib():
    if not os.path.exists(sister_amd.LIB_PATH):
        from sister_amd import build
        build.build()
    return ctypes.CDLL(sister_amd.LIB_PATH)


def test_library_exports_every_declared_symbol(hiplib):
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(hiplib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    assert set(_declared_functions()) == set(sister_amd._SIGS)


def test_workspace_and_domain_queries(hiplib):
    # pure host arithmetic: no device needed
    assert sister_amd.workspace_bytes(640, 480, 64) > 0
    assert sister_amd.workspace_bytes(642, 480, 64) == 0      # W not a multiple of 4
    assert sister_amd.workspace_bytes(640, 480, 60) == 0      # D not a multiple of 8
    assert sister_amd.workspace_bytes(640, 480, 520) == 0     # D above 512
    assert not sister_amd.outside_reference_domain(256)
    assert sister_amd.outside_reference_domain(384)


def test_context_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(sister_amd.SisterError):
        sister_amd.Context()


def test_bgr_to_gray_identity_on_gray_replicated():
    g = np.random.default_rng(0).integers(0, 256, size=(8, 12)).astype(np.uint8)
    bgr = np.repeat(g[..., None], 3, axis=2)
    assert np.array_equal(sister_amd.bgr_to_gray(bgr), g)
    # weights of cvtColor(BGR2GRAY): pure blue / green / red
    px = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255]]], np.uint8)
    assert sister_amd.bgr_to_gray(px).tolist() == [[29, 150, 76]]


def test_synthetic_rigs_deterministic():
    a = synth.rig_small(64, 48, 16, 5)
    b = synth.rig_small(64, 48, 16, 5)
    for x, y in zip(a["views"], b["views"]):
        assert np.array_equal(x, y)
    assert synth.batch_shapes(32, 1000) == synth.batch_shapes(32, 1000)
    assert all(W % 4 == 0 and H % 4 == 0 and D % 8 == 0 for W, H, D in synth.batch_shapes(32, 1000))


def test_shift_convention():
    # right(x, y) = C(x + d, y): a constant-d rig is an exact shift away from the borders
    rng = np.random.default_rng(1)
    C = rng.integers(0, 256, size=(20, 40)).astype(np.float64)
    gt = np.full((20, 40), 5.0)
    views = synth.make_views(C, gt, rng, noise_sigma=0.0)
    center, right, top, left, bottom = views
    assert np.array_equal(right[:, :30], center[:, 5:35])
    assert np.array_equal(left[:, 10:], center[:, 5:35])
    assert np.array_equal(top[10:], center[5:15])
    assert np.array_equal(bottom[:10], center[5:15])
