"""Pins of the helper-class glue (CPU only).

The oracle's orchestration (oracle/orchestrate.c) restates the glue of
SisterMultiviewDisparities.hpp -- padding, the four frames built by flip and
transpose, the per-frame confidence masks and their map back to original
coordinates, the realignment and mask-weighted sum, the crop and the x255
scaling.  The reference-kernel build shares that file, so the kernel pins do
not cover it.  Here a second restatement written from the header with numpy
(tests/glue_np.py) is compared with it on every intermediate, driving the
reference's own compiled kernels where they are built; planted glue mistakes
must change the outputs; and a rig whose true disparity is known must come
out at that disparity.
"""
import numpy as np
import pytest

from sister_amd import synth
from tests import glue_np

RIGS = [(64, 48, 16, 51), (96, 64, 24, 52), (80, 60, 32, 53)]
NAMES = ("multiview", "horizontal", "vertical")


def _kernel_sets(oracle_mod):
    sets = [("oracle kernels", oracle_mod.kernels(False))]
    if oracle_mod.ref_available():
        sets.append(("reference kernels", oracle_mod.kernels(True)))
    return sets


@pytest.mark.parametrize("W,H,D,seed", RIGS)
def test_glue_restatement_matches_oracle(oracle_mod, W, H, D, seed):
    rig = synth.rig_small(W, H, D, seed)
    want = oracle_mod.compute(rig["views"], D, modes=7, debug=True)
    for name, K in _kernel_sets(oracle_mod):
        tr = {}
        got = glue_np.compute_disparities(K, rig["views"], D, modes=7, trace=tr)
        for k in NAMES:
            assert np.array_equal(got[k], want[k]), (name, k)
        # frame images: their census codes are the oracle's per-frame codes
        for F, (imc, imn) in enumerate(tr["frames"]):
            assert np.array_equal(K.census(imc), want["census"][2 * F]), (name, "census centre", F)
            assert np.array_equal(K.census(imn), want["census"][2 * F + 1]), (name, "census neighbour", F)
        for F in range(4):
            for nm in ("wta_l", "wta_r", "med_l", "med_r"):
                assert np.array_equal(tr[f"{nm}_{F}"], want[nm][F]), (name, nm, F)
        # masks in original coordinates: all four in mode 0; 0/180 in mode 1; 90/270 in mode 2
        for F in range(4):
            assert np.array_equal(tr["mask_0"][F], want["mask"][F]), (name, "mask", F)
        for F in (0, 2):
            assert np.array_equal(tr["mask_1"][F], want["mask"][F]), (name, "mask mode 1", F)
        for F in (1, 3):
            assert np.array_equal(tr["mask_2"][F], want["mask"][F]), (name, "mask mode 2", F)
        for m in range(3):
            assert np.array_equal(tr[f"msum_{m}"], want["msum"][m]), (name, "msum", m)
            assert np.array_equal(tr[f"sum_{m}"], want["sum"][m]), (name, "sum", m)
            assert np.array_equal(tr[f"disp_pad_{m}"], want["disp_pad"][m]), (name, "disp_pad", m)


@pytest.mark.parametrize("mutation", glue_np.MUTATIONS)
def test_planted_glue_mistakes_change_the_outputs(oracle_mod, mutation):
    """Each planted mistake (frames 90/270 or 0/180 swapped, flip code 1 written as 0,
    a mask not mapped back, a wrong flip of a mask or a volume, a crop off by
    one row) must change at least one output map, so the comparison above
    would catch it."""
    W, H, D, seed = RIGS[1]
    rig = synth.rig_small(W, H, D, seed)
    want = oracle_mod.compute(rig["views"], D, modes=7)
    K = oracle_mod.kernels(False)
    got = glue_np.compute_disparities(K, rig["views"], D, modes=7, mutate=mutation)
    assert any(not np.array_equal(got[k], want[k]) for k in NAMES), mutation


def _interior(W, H, d):
    m = d + 8
    return (slice(m, H - m), slice(m, W - m))


@pytest.mark.parametrize("W,H,D,d,seed", [(128, 96, 32, 9, 61), (112, 104, 24, 17, 62)])
def test_constant_disparity_is_recovered(oracle_mod, W, H, D, d, seed):
    """Ground truth: views shifted by exactly d pixels with no noise.  The
    multiview, horizontal and vertical maps must read d at 99% or more of the
    pixels at least d + 8 from every border, through the oracle and through
    the numpy glue over the reference's kernels.  A glue mistake that pairs a
    view with the wrong frame or mirror sends its matches away from d."""
    rig = synth.rig_shift(W, H, D, d, seed)
    outs = [oracle_mod.compute(rig["views"], D, modes=7)]
    if oracle_mod.ref_available():
        outs.append(glue_np.compute_disparities(oracle_mod.kernels(True), rig["views"], D, modes=7))
    ys, xs = _interior(W, H, d)
    for out in outs:
        for k in NAMES:
            v = out[k][ys, xs]
            frac = float(np.mean(v == 255 * d))
            assert frac >= 0.99, (k, frac)


def test_constant_disparity_rejects_swapped_frames(oracle_mod):
    """The same ground truth check fails for a glue that swaps the top and
    bottom views, or the right and left views."""
    W, H, D, d, seed = 128, 96, 32, 9, 61
    rig = synth.rig_shift(W, H, D, d, seed)
    K = oracle_mod.kernels(False)
    ys, xs = _interior(W, H, d)
    for mutation, key in (("swap_90_270", "vertical"), ("swap_0_180", "horizontal")):
        out = glue_np.compute_disparities(K, rig["views"], D, modes=7, mutate=mutation)
        frac = float(np.mean(out[key][ys, xs] == 255 * d))
        assert frac < 0.9, (mutation, frac)
