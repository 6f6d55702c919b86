"""GPU: batched plans, graph replay, serving, error paths, full-size parity.

Every result is compared bit-exactly with the oracle (small rigs) or with the
reference's own kernels (oracle/_ref, the 1920x1080 D=192 rig), or with the
same rig computed alone (size-independent properties).
"""
import os
import time

import numpy as np
import pytest

import sister_amd
from sister_amd import synth

pytestmark = pytest.mark.gpu

MIX = [(64, 48, 16, 51), (96, 64, 24, 52), (80, 60, 72, 53), (128, 96, 32, 54), (64, 48, 16, 55), (72, 40, 64, 56)]


def _device_rig(ctx, rig, torch):
    H, W = rig["views"][0].shape
    views = [torch.from_numpy(v).cuda() for v in rig["views"]]
    outs = {"multiview": torch.zeros((H, W), dtype=torch.uint16, device="cuda"),
            "disp": torch.zeros((H, W), dtype=torch.uint16, device="cuda"),
            "subpix": torch.zeros((H, W), dtype=torch.float32, device="cuda")}
    return dict(views=views, outs=outs, struct=ctx.rig_struct(views, rig["D"], outs))


def test_batched_mixed_sizes_match_oracle(gpu_ctx, oracle_mod):
    import torch
    rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in MIX]
    dev = [_device_rig(gpu_ctx, r, torch) for r in rigs]
    print("[batch] 6 mixed rigs in one plan", flush=True)
    gpu_ctx.run_device([d["struct"] for d in dev], sister_amd.MODE_MULTIVIEW)
    gpu_ctx.sync()
    for r, d in zip(rigs, dev):
        ref = oracle_mod.compute(r["views"], r["D"], modes=1)
        assert np.array_equal(d["outs"]["multiview"].cpu().numpy(), ref["multiview"])
        assert np.array_equal(d["outs"]["disp"].cpu().numpy(), ref["disp"])
        assert float(np.abs(d["outs"]["subpix"].cpu().numpy() - ref["subpix"]).max()) <= 1e-3


def test_batched_fused_vertical_mixed_sizes_match_oracle(gpu_ctx, oracle_mod):
    """The fused vertical sweep (forced on) over rigs of one D with different
    widths and heights in one plan: per-rig block ranges, rigs that end
    segments early (H' of 104, 136, 168 and 92 rows: 6 to 11 segments of 16),
    plus a second D group beside them."""
    import torch
    specs = [(200, 40, 32, 61), (96, 72, 32, 62), (64, 104, 32, 63), (152, 28, 32, 64), (80, 60, 24, 65)]
    rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in specs]
    dev = [_device_rig(gpu_ctx, r, torch) for r in rigs]
    print("[batch] 5 mixed rigs, fused vertical sweep", flush=True)
    gpu_ctx.vsweep(1)
    try:
        gpu_ctx.run_device([d["struct"] for d in dev], sister_amd.MODE_MULTIVIEW)
        gpu_ctx.sync()
    finally:
        gpu_ctx.vsweep(-1)
    for r, d in zip(rigs, dev):
        ref = oracle_mod.compute(r["views"], r["D"], modes=1)
        assert np.array_equal(d["outs"]["multiview"].cpu().numpy(), ref["multiview"])
        assert np.array_equal(d["outs"]["disp"].cpu().numpy(), ref["disp"])
        assert float(np.abs(d["outs"]["subpix"].cpu().numpy() - ref["subpix"]).max()) <= 1e-3


def test_batched_eight_rigs_medians_match_oracle(gpu_ctx, oracle_mod):
    """Eight rigs give 64 (rig, map) blocks: enough for the serial median to
    sort each position's six fixed values itself (one rig takes the prepass
    form, which the stage tests cover).  Every median map, then the outputs."""
    import torch
    specs = MIX + [(64, 40, 16, 57), (48, 48, 24, 58)]
    rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in specs]
    dev = [_device_rig(gpu_ctx, r, torch) for r in rigs]
    print("[batch] 8 rigs in one plan, medians", flush=True)
    gpu_ctx.run_device([d["struct"] for d in dev], sister_amd.MODE_MULTIVIEW)
    gpu_ctx.sync()
    for k, ((W, H, D, _), r, d) in enumerate(zip(specs, rigs, dev)):
        ref = oracle_mod.compute(r["views"], D, modes=1, debug=True)
        Wp, Hp = W + 2 * D, H + 2 * D
        shapes = [(Hp, Wp), (Wp, Hp), (Hp, Wp), (Wp, Hp)]
        for F in range(4):
            for side, nm in ((0, "med_l"), (1, "med_r")):
                g = gpu_ctx.fetch(k, "med", 2 * F + side, shapes[F], np.uint16)
                assert np.array_equal(g, ref[nm][F].astype(np.uint16)), (k, F, side)
        assert np.array_equal(d["outs"]["multiview"].cpu().numpy(), ref["multiview"])
        assert np.array_equal(d["outs"]["disp"].cpu().numpy(), ref["disp"])


def test_graph_replays_match_eager(gpu_ctx):
    import torch
    rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in MIX[:4]]
    dev = [_device_rig(gpu_ctx, r, torch) for r in rigs]
    gpu_ctx.run_device([d["struct"] for d in dev], 1)
    gpu_ctx.sync()
    eager = [d["outs"]["disp"].cpu().numpy().copy() for d in dev]
    other = _device_rig(gpu_ctx, synth.rig_small(64, 48, 16, 99), torch)
    print("[graph] capture 4 rigs, replay 3x with eager work between", flush=True)
    gpu_ctx.graph_capture([d["struct"] for d in dev], 1)
    for rep in range(3):
        for d in dev:
            d["outs"]["disp"].zero_()
        torch.cuda.synchronize()
        gpu_ctx.run_device([other["struct"]], 1)     # unrelated eager work in between
        gpu_ctx.sync()
        gpu_ctx.graph_launch()
        gpu_ctx.sync()
        for e, d in zip(eager, dev):
            assert np.array_equal(d["outs"]["disp"].cpu().numpy(), e), rep


def test_deterministic_and_plain_equals_batched(gpu_ctx):
    import torch
    rig = synth.rig_small(160, 120, 40, 61)
    a = _device_rig(gpu_ctx, rig, torch)
    gpu_ctx.run_device([a["struct"]], 1)
    gpu_ctx.sync()
    first = a["outs"]["disp"].cpu().numpy().copy()
    gpu_ctx.run_device([a["struct"]], 1)
    gpu_ctx.sync()
    assert np.array_equal(a["outs"]["disp"].cpu().numpy(), first)
    # the same rig inside a larger batch (different launch geometry)
    others = [_device_rig(gpu_ctx, synth.rig_small(W, H, D, s), torch) for W, H, D, s in MIX]
    b = _device_rig(gpu_ctx, rig, torch)
    gpu_ctx.run_device([o["struct"] for o in others] + [b["struct"]], 1)
    gpu_ctx.sync()
    assert np.array_equal(b["outs"]["disp"].cpu().numpy(), first)


def _host_rig_struct(r, outs):
    s = sister_amd.RigStruct()
    s.width, s.height, s.stride, s.disp_count = r["W"], r["H"], r["W"], r["D"]
    for k in range(5):
        s.view[k] = r["views"][k].ctypes.data
    s.out_multiview = outs["multiview"].ctypes.data
    s.out_disp = outs["disp"].ctypes.data
    return s


def test_serving_alone_equals_mix():
    ctx = sister_amd.Context()
    try:
        ctx.serve_start(3)
        rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in MIX]
        keep = []

        def submit(r):
            o = {"multiview": np.zeros((r["H"], r["W"]), np.uint16), "disp": np.zeros((r["H"], r["W"]), np.uint16)}
            s = _host_rig_struct(r, o)
            keep.append((s, o, r))
            return ctx.submit(s, 1), o

        print("[serve] rig alone", flush=True)
        t, alone = submit(rigs[0])
        ctx.wait(t)
        print("[serve] same rig in a mix of 12 arrivals", flush=True)
        pending = []
        for i in range(12):
            pending.append(submit(rigs[i % len(rigs)]))
            time.sleep(0.001)
        for t, o in pending:
            ctx.wait(t)
        for i, (t, o) in enumerate(pending):
            if i % len(rigs) == 0:
                assert np.array_equal(o["disp"], alone["disp"])
                assert np.array_equal(o["multiview"], alone["multiview"])
        # every rig of the mix equals its own single run
        for i, r in enumerate(rigs):
            single = ctx.compute_rig(r["views"], r["D"], modes=1, subpix=False)
            assert np.array_equal(pending[i][1]["disp"], single["disp"]), i
        st = ctx.stats()
        assert st["completed"] == 13 and st["rejected"] == 0
        assert st["p99_ms"] >= st["p50_ms"] > 0
        ctx.serve_stop()
    finally:
        ctx.close()


def test_serving_fused_vertical_matches_oracle(oracle_mod):
    """The serving loop's own plans on the fused vertical sweep (forced on for
    the context): rigs of several widths and heights arriving together, each
    result equal to the oracle."""
    specs = [(200, 40, 32, 81), (96, 72, 32, 82), (64, 104, 32, 83), (152, 28, 32, 84), (80, 60, 24, 85)]
    rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in specs]
    ctx = sister_amd.Context()
    try:
        ctx.vsweep(1)
        ctx.serve_start(3)
        keep, pending = [], []
        print("[serve] 5 arrivals on the fused vertical sweep", flush=True)
        for r in rigs:
            o = {"multiview": np.zeros((r["H"], r["W"]), np.uint16), "disp": np.zeros((r["H"], r["W"]), np.uint16)}
            st = _host_rig_struct(r, o)
            keep.append(st)
            pending.append((ctx.submit(st, 1), o, r))
        for t, o, r in pending:
            ctx.wait(t)
        for t, o, r in pending:
            ref = oracle_mod.compute(r["views"], r["D"], modes=1, subpix=False)
            assert np.array_equal(o["disp"], ref["disp"])
            assert np.array_equal(o["multiview"], ref["multiview"])
        ctx.serve_stop()
    finally:
        ctx.close()


def test_serving_device_buffers_equal_run_device(gpu_ctx):
    """Device-resident submissions: computed in place, equal to the same rigs
    run directly."""
    import torch
    rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in MIX]
    ref = [_device_rig(gpu_ctx, r, torch) for r in rigs]
    gpu_ctx.run_device([d["struct"] for d in ref], 1)
    gpu_ctx.sync()
    ctx = sister_amd.Context()
    try:
        ctx.serve_start(3)
        dev = [_device_rig(ctx, r, torch) for r in rigs for _ in range(2)]
        torch.cuda.synchronize()
        print("[serve-device] 12 device-resident requests", flush=True)
        tickets = [ctx.submit(d["struct"], 1, device=True) for d in dev]
        for t in tickets:
            ctx.wait(t)
        for i, d in enumerate(dev):
            assert torch.equal(d["outs"]["disp"], ref[i // 2]["outs"]["disp"]), i
            assert torch.equal(d["outs"]["multiview"], ref[i // 2]["outs"]["multiview"]), i
        st = ctx.stats()
        assert st["completed"] == 12 and st["rejected"] == 0
        ctx.serve_stop()
    finally:
        ctx.close()


def test_serving_admission_by_budget():
    small = synth.rig_small(64, 48, 16, 71)
    big = synth.rig_small(320, 240, 64, 72)
    fp_small = sister_amd.workspace_bytes(64, 48, 16)
    ctx = sister_amd.Context(hbm_budget=int(fp_small * 3.5))
    try:
        ctx.serve_start(2)
        o = {"multiview": np.zeros((240, 320), np.uint16), "disp": np.zeros((240, 320), np.uint16)}
        with pytest.raises(sister_amd.SisterError) as e:
            ctx.submit(_host_rig_struct(big, o), 1)
        assert e.value.code == sister_amd.SISTER_EBUSY
        outs = []
        tickets = []
        for i in range(8):   # more than fit at once: admitted as memory frees
            oo = {"multiview": np.zeros((48, 64), np.uint16), "disp": np.zeros((48, 64), np.uint16)}
            s = _host_rig_struct(small, oo)
            outs.append((s, oo))
            tickets.append(ctx.submit(s, 1))
        for t in tickets:
            ctx.wait(t)
        first = outs[0][1]["disp"]
        assert all(np.array_equal(oo["disp"], first) for _, oo in outs)
        st = ctx.stats()
        assert st["completed"] == 8 and st["rejected"] == 1
        assert st["hbm_peak"] <= int(fp_small * 3.5)
        ctx.serve_stop()
    finally:
        ctx.close()


@pytest.mark.parametrize("W,H,D", [(64, 48, 60), (66, 48, 16), (64, 46, 16), (64, 48, 0), (64, 48, 520), (16376, 48, 8)])
def test_invalid_rigs_rejected(gpu_ctx, W, H, D):
    views = [np.zeros((H, W), np.uint8)] * 5
    with pytest.raises(sister_amd.SisterError) as e:
        gpu_ctx.compute_rig(views, D, modes=1)
    assert e.value.code == sister_amd.SISTER_EINVAL


def test_full_1080_d192_bit_exact_vs_reference_kernels(gpu_ctx, oracle_mod):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built")
    rig = synth.rig_plain_1080()
    print("[full] 1920x1080 D=192 seed=2 vs reference kernels", flush=True)
    got = gpu_ctx.compute_rig(rig["views"], rig["D"], modes=1)
    ref = oracle_mod.compute(rig["views"], rig["D"], modes=1, use_ref=True, subpix=False)
    assert np.array_equal(got["disp"], ref["disp"])
    assert np.array_equal(got["multiview"], ref["multiview"])
    # size-independent properties of the full-size output
    assert got["disp"].max() <= rig["D"] - 1
    assert np.array_equal(got["multiview"].astype(np.uint32), np.minimum(255 * got["disp"].astype(np.uint32), 65535))
    assert np.all(np.abs(got["subpix"] - got["disp"]) <= 0.5 + 1e-6)


def test_graph_replays_fused_vertical_match_oracle(gpu_ctx, oracle_mod):
    """A captured batch on the fused vertical sweep (forced on at capture), in
    two graph parts, replayed three times with unrelated eager work between
    replays: every replay equals the oracle."""
    import torch
    specs = [(200, 40, 32, 71), (96, 72, 32, 72), (64, 104, 32, 73), (152, 28, 32, 74)]
    rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in specs]
    refs = [oracle_mod.compute(r["views"], r["D"], modes=1) for r in rigs]
    dev = [_device_rig(gpu_ctx, r, torch) for r in rigs]
    other = _device_rig(gpu_ctx, synth.rig_small(64, 48, 16, 75), torch)
    print("[graph] capture 4 rigs on the fused sweep, replay 3x", flush=True)
    gpu_ctx.vsweep(1)
    try:
        gpu_ctx.graph_capture([d["struct"] for d in dev], sister_amd.MODE_MULTIVIEW)
    finally:
        gpu_ctx.vsweep(-1)
    for rep in range(3):
        for d in dev:
            d["outs"]["disp"].zero_()
            d["outs"]["multiview"].zero_()
        torch.cuda.synchronize()
        gpu_ctx.run_device([other["struct"]], 1)
        gpu_ctx.sync()
        gpu_ctx.graph_launch()
        gpu_ctx.sync()
        for ref, d in zip(refs, dev):
            assert np.array_equal(d["outs"]["disp"].cpu().numpy(), ref["disp"]), rep
            assert np.array_equal(d["outs"]["multiview"].cpu().numpy(), ref["multiview"]), rep


def test_full_1080_d192_fused_vertical_equals_separate(gpu_ctx):
    """At full size the fused vertical sweep (forced on for the one rig: 16
    blocks of 144 columns, 92 segments of 16 rows a pass) gives what the
    separate families give; those equal the reference kernels at this size
    (the test above).  Every output, S-derived sub-pixel map included."""
    rig = synth.rig_plain_1080()
    print("[full] 1920x1080 D=192 seed=2 fused vs separate vertical families", flush=True)
    gpu_ctx.vsweep(0)
    try:
        sep = gpu_ctx.compute_rig(rig["views"], rig["D"], modes=7)
        gpu_ctx.vsweep(1)
        fused = gpu_ctx.compute_rig(rig["views"], rig["D"], modes=7)
        gpu_ctx.sweep_in_place(True)   # partial sums over the rows family's bytes
        in_place = gpu_ctx.compute_rig(rig["views"], rig["D"], modes=7)
    finally:
        gpu_ctx.vsweep(-1)
        gpu_ctx.sweep_in_place(False)
    for k in ("multiview", "horizontal", "vertical", "disp", "subpix"):
        assert np.array_equal(fused[k], sep[k]), (k, int((fused[k] != sep[k]).sum()))
        assert np.array_equal(in_place[k], sep[k]), (k, int((in_place[k] != sep[k]).sum()))


def test_serving_caller_errors_while_serving():
    """Caller threads fail calls (invalid submissions, unknown tickets) and
    read the error text while the serving thread builds and retires plans:
    each caller reads its own failure's text, every request completes, and
    results equal the same rigs alone."""
    import threading
    import ctypes as C
    rigs = [synth.rig_small(W, H, D, s) for W, H, D, s in MIX]
    ctx = sister_amd.Context()
    try:
        singles = [ctx.compute_rig(r["views"], r["D"], modes=1, subpix=False) for r in rigs]
        ctx.serve_start(3)
        L = sister_amd.lib()
        errors = []

        def hammer(seed):
            bad = sister_amd.RigStruct()
            bad.width, bad.height, bad.stride, bad.disp_count = 66, 48, 66, 16   # W not a multiple of 4
            t = C.c_uint64()
            for i in range(3000):
                if i % 2:
                    rc = L.sister_submit(ctx._h, C.byref(bad), 1, C.byref(t))
                    want, text = sister_amd.SISTER_EINVAL, b"invalid rig shape"
                else:
                    d = C.c_int()
                    rc = L.sister_poll(ctx._h, 10**12 + seed * 10**6 + i, C.byref(d))
                    want, text = sister_amd.SISTER_ENOTFOUND, b"unknown ticket"
                msg = L.sister_last_error(ctx._h)
                if rc != want or not msg.startswith(text):
                    errors.append((seed, i, rc, msg))
                    return

        threads = [threading.Thread(target=hammer, args=(k,)) for k in range(2)]
        for th in threads:
            th.start()
        print("[serve] 24 requests while two threads fail calls", flush=True)
        pending = []
        for i in range(24):
            r = rigs[i % len(rigs)]
            o = {"multiview": np.zeros((r["H"], r["W"]), np.uint16), "disp": np.zeros((r["H"], r["W"]), np.uint16)}
            s = _host_rig_struct(r, o)
            pending.append((ctx.submit(s, 1), o, s, i % len(rigs)))
        for t, o, s, k in pending:
            ctx.wait(t)
        for th in threads:
            th.join()
        assert not errors, errors[:3]
        for t, o, s, k in pending:
            assert np.array_equal(o["disp"], singles[k]["disp"]), k
        st = ctx.stats()
        assert st["completed"] == 24
        ctx.serve_stop()
    finally:
        ctx.close()
