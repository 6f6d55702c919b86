"""GPU: the dataset harness end to end on a synthetic dataset in the
documented layout (<object>/<distance>/<baseline>mm/<view>.png): every rig's
16-bit maps equal the oracle's."""
import numpy as np
import pytest

from sister_amd import dataset, imageio, synth

pytestmark = pytest.mark.gpu


def test_run_over_exported_layout(tmp_path, oracle_mod):
    rigs = {("washer", "5cm", "010mm"): synth.rig_small(96, 64, 24, 91),
            ("washer", "5cm", "025mm"): synth.rig_small(96, 64, 24, 92),
            ("hexa_nut", "1cm", "100mm"): synth.rig_small(64, 48, 24, 93)}
    for (o, d, b), r in rigs.items():
        f = tmp_path / o / d / b
        f.mkdir(parents=True)
        for i, (v, img) in enumerate(zip(dataset.VIEWS, r["views"])):
            # half the views as BGR images with the gray value in every channel
            imageio.write_png(f / f"{v}.png", np.repeat(img[..., None], 3, 2) if i % 2 else img)
    print("[dataset] 3 rigs through dataset.run", flush=True)
    done = dataset.run(tmp_path, 24, modes=7)
    assert sorted(x[:3] for x in done) == sorted(rigs)
    for (o, d, b), r in rigs.items():
        ref = oracle_mod.compute(r["views"], 24, modes=7)
        for k in ("multiview", "horizontal", "vertical"):
            got = imageio.read_png(tmp_path / o / d / b / f"disp_{k}.png")
            assert got.dtype == np.uint16 and np.array_equal(got, ref[k]), (o, d, b, k)
