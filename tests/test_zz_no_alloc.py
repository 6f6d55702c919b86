"""GPU: the allocation-free entry point (sister_run_device_noalloc) against
sister_run_device on the same rigs, called through the C ABI: two calls in a
row, the second on a smaller rig, every output map identical; a rig beyond the
scratch array's size is refused before anything is launched."""
import ctypes as C

import numpy as np
import pytest

import sister_amd
from sister_amd import synth

pytestmark = pytest.mark.gpu

OUTS = ("multiview", "horizontal", "vertical", "disp", "subpix")


def _device_rig(ctx, rig, torch):
    H, W = rig["views"][0].shape
    views = [torch.from_numpy(v).cuda() for v in rig["views"]]
    o = {k: torch.full((H, W), 7, dtype=torch.float32 if k == "subpix" else torch.uint16, device="cuda")
         for k in OUTS}
    return dict(views=views, outs=o, struct=ctx.rig_struct(views, rig["D"], o))


def test_noalloc_equals_run_device():
    import torch
    L = sister_amd.lib()
    ctx = sister_amd.Context()
    try:
        for W, H, D, seed in ((640, 480, 64, 91), (320, 240, 32, 92)):
            rig = synth.rig_scene(W, H, D, seed)
            want = _device_rig(ctx, rig, torch)
            got = _device_rig(ctx, rig, torch)
            torch.cuda.synchronize()
            print(f"[noalloc] {W}x{H} D={D}: run_device, then run_device_noalloc", flush=True)
            ctx.run_device([want["struct"]], sister_amd.MODE_ALL)
            ctx.sync()
            rc = L.sister_run_device_noalloc(ctx._h, C.byref(got["struct"]), 1, sister_amd.MODE_ALL)
            assert rc == sister_amd.SISTER_OK, L.sister_last_error(ctx._h)
            ctx.sync()
            for k in OUTS:
                assert torch.equal(got["outs"][k], want["outs"][k]), (W, H, D, k)
        # beyond the scratch array: refused, nothing launched
        big = synth.rig_scene(1280, 720, 64, 93)
        dev = _device_rig(ctx, big, torch)
        assert sister_amd.workspace_bytes(1280, 720, 64) > sister_amd.NOALLOC_MAX_BYTES
        rc = L.sister_run_device_noalloc(ctx._h, C.byref(dev["struct"]), 1, sister_amd.MODE_ALL)
        assert rc == sister_amd.SISTER_ENOMEM
        assert b"scratch" in L.sister_last_error(ctx._h)
        assert torch.all(dev["outs"]["multiview"] == 7)
    finally:
        ctx.close()
