"""GPU: the autotuner.  Tuned forms change launch shapes only: outputs must be
identical to the default forms', eager and captured, and the forms file must
read back into a new context without timing anything again."""
import os

import numpy as np
import pytest

import sister_amd
from sister_amd import synth

pytestmark = pytest.mark.gpu

MIX = [(64, 48, 16, 51), (96, 64, 24, 52), (80, 60, 72, 53), (128, 96, 32, 54), (64, 48, 16, 55),
       (72, 40, 64, 56), (200, 40, 32, 57), (96, 72, 32, 58)]


def _device_rig(ctx, rig, torch):
    H, W = rig["views"][0].shape
    views = [torch.from_numpy(v).cuda() for v in rig["views"]]
    outs = {"multiview": torch.zeros((H, W), dtype=torch.uint16, device="cuda"),
            "disp": torch.zeros((H, W), dtype=torch.uint16, device="cuda"),
            "subpix": torch.zeros((H, W), dtype=torch.float32, device="cuda")}
    return dict(views=views, outs=outs, struct=ctx.rig_struct(views, rig["D"], outs))


def _snap(dev):
    return [{k: v.clone() for k, v in d["outs"].items()} for d in dev]


def _same(dev, ref):
    return [i for i, (d, r) in enumerate(zip(dev, ref)) if not all(np.array_equal(
        d["outs"][k].cpu().numpy(), r[k].cpu().numpy()) for k in r)]


def test_tuned_forms_give_identical_outputs(tmp_path, oracle_mod):
    import torch
    rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in MIX]
    path = str(tmp_path / "forms.txt")
    ctx = sister_amd.Context()
    try:
        dev = [_device_rig(ctx, r, torch) for r in rigs]
        structs = [d["struct"] for d in dev]
        print("[tune] default forms, eager", flush=True)
        ctx.run_device(structs, 1)
        ctx.sync()
        ref = _snap(dev)
        for r, o in zip(rigs[:3], ref[:3]):
            want = oracle_mod.compute(r["views"], r["D"], modes=1)
            assert np.array_equal(o["disp"].cpu().numpy(), want["disp"])
        print("[tune] autotune (graph captures of every candidate form)", flush=True)
        ctx.autotune(structs, 1, path)
        table = ctx.tune_table()
        print(table, flush=True)
        assert "batch 8 " in table and table.count("group ") >= 1
        assert os.path.exists(path)
        for d in dev:
            for v in d["outs"].values():
                v.zero_()
        torch.cuda.synchronize()
        print("[tune] tuned graph replay", flush=True)
        ctx.graph_launch()
        ctx.sync()
        assert not _same(dev, ref)
        for d in dev:
            for v in d["outs"].values():
                v.zero_()
        torch.cuda.synchronize()
        print("[tune] tuned forms, eager", flush=True)
        ctx.run_device(structs, 1)
        ctx.sync()
        assert not _same(dev, ref)
    finally:
        ctx.close()
    # a new context reads the file: same table, nothing timed again
    ctx2 = sister_amd.Context()
    try:
        dev2 = [_device_rig(ctx2, r, torch) for r in rigs]
        ctx2.autotune([d["struct"] for d in dev2], 1, path)
        assert ctx2.tune_table() == table
        ctx2.graph_launch()
        ctx2.sync()
        assert not _same(dev2, ref)
    finally:
        ctx2.close()
