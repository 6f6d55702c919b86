"""GPU parity: the HIP path against the oracle, stage by stage, bit-exact.

Integer results (census codes, WTA maps, medians, masks, weighted cost sum,
SGM aggregate S, disparities, API maps) must be identical; the sub-pixel map
(build-defined parabola, no reference counterpart) within 1e-3 px.
"""
import numpy as np
import pytest

from sister_amd import synth

pytestmark = pytest.mark.gpu

SMALL = [  # (W, H, D, seed): D covers one partial lane slot, one full, several
    (64, 48, 16, 3),
    (96, 64, 24, 4),
    (128, 96, 32, 5),
    (72, 40, 64, 6),
    (80, 60, 72, 7),
    (160, 120, 40, 8),
    (40, 36, 8, 9),
]


def _compare_stages(ctx, O, W, H, D, seed, modes=7):
    rig = synth.rig_small(W, H, D, seed)
    ref = O.compute(rig["views"], D, modes=modes, debug=True)
    ctx.keep_sum(True)
    try:
        got = ctx.compute_rig(rig["views"], D, modes=modes)
    finally:
        ctx.keep_sum(False)
    Wp, Hp = W + 2 * D, H + 2 * D
    shapes = [(Hp, Wp), (Wp, Hp), (Hp, Wp), (Wp, Hp)]
    bad = []
    for F in range(4):
        for img in range(2):
            g = ctx.fetch(0, "census", 2 * F + img, shapes[F], np.uint64)
            if not np.array_equal(g, ref["census"][2 * F + img]):
                bad.append(("census", F, img, int((g != ref["census"][2 * F + img]).sum())))
        for side, nm in ((0, "wta_l"), (1, "wta_r")):
            g = ctx.fetch(0, "wta", 2 * F + side, shapes[F], np.uint16)
            if not np.array_equal(g, ref[nm][F].astype(np.uint16)):
                bad.append(("wta", F, side, int((g != ref[nm][F]).sum())))
        for side, nm in ((0, "med_l"), (1, "med_r")):
            g = ctx.fetch(0, "med", 2 * F + side, shapes[F], np.uint16)
            if not np.array_equal(g, ref[nm][F].astype(np.uint16)):
                bad.append(("med", F, side, int((g != ref[nm][F]).sum())))
        g = ctx.fetch(0, "mask", F, (Hp, Wp), np.uint8)
        if not np.array_equal(g, ref["mask"][F]):
            bad.append(("mask", F, int((g != ref["mask"][F]).sum())))
    # the workspace holds the last mode's cost sum and aggregate
    last = max(m for m in range(3) if modes & (1 << m))
    Dp = (D + 15) // 16 * 16
    g = ctx.fetch(0, "msum", 0, (Hp, Wp, Dp), np.uint8)[:, :, :D]
    if not np.array_equal(g.astype(np.uint16), ref["msum"][last]):
        bad.append(("msum", int((g != ref["msum"][last]).sum())))
    g = ctx.fetch(0, "sum", 0, (Hp, Wp, Dp), np.uint16)[:, :, :D]
    if not np.array_equal(g, ref["sum"][last]):
        diff = np.argwhere(g != ref["sum"][last])
        bad.append(("sum", len(diff), diff[:3].tolist()))
    for k in ("multiview", "horizontal", "vertical", "disp"):
        if k in ref and k in got and not np.array_equal(got[k], ref[k]):
            bad.append((k, int((got[k] != ref[k]).sum())))
    if "subpix" in ref:
        err = float(np.abs(got["subpix"] - ref["subpix"]).max())
        if err > 1e-3:
            bad.append(("subpix", err))
    return bad


@pytest.mark.parametrize("W,H,D,seed", SMALL)
def test_stages_bit_exact(gpu_ctx, oracle_mod, W, H, D, seed):
    print(f"[parity] W={W} H={H} D={D} seed={seed}", flush=True)
    gpu_ctx.vsweep(0)
    try:
        bad = _compare_stages(gpu_ctx, oracle_mod, W, H, D, seed)
    finally:
        gpu_ctx.vsweep(-1)
    assert not bad, bad


# the fused vertical sweep (forced on at sizes where it would not be chosen):
# one block (144 columns) and several, segment boundaries (16 rows) inside and
# at the end, partial last blocks, every disparity layout up to 192
FUSED = [
    (64, 48, 16, 3),
    (96, 64, 24, 4),
    (128, 96, 32, 5),
    (72, 40, 64, 6),
    (80, 60, 72, 7),
    (160, 120, 40, 8),
    (40, 36, 8, 9),
    (200, 44, 56, 10),
    (96, 52, 120, 11),
    (64, 40, 192, 12),
]


@pytest.mark.parametrize("W,H,D,seed", FUSED)
def test_stages_bit_exact_fused_vertical(gpu_ctx, oracle_mod, W, H, D, seed):
    print(f"[parity fused] W={W} H={H} D={D} seed={seed}", flush=True)
    gpu_ctx.vsweep(1)
    try:
        bad = _compare_stages(gpu_ctx, oracle_mod, W, H, D, seed)
    finally:
        gpu_ctx.vsweep(-1)
    assert not bad, bad


# the fused sweep's full-lane form (D a multiple of 64, no debug sums: cost
# rows, rows-family bytes and partial sums in row planes): outputs only
@pytest.mark.parametrize("W,H,D,seed", [(96, 64, 64, 13), (64, 48, 128, 14), (64, 40, 192, 15), (400, 40, 64, 16)])
def test_fused_vertical_full_lanes_outputs(gpu_ctx, oracle_mod, W, H, D, seed):
    print(f"[parity fused full] W={W} H={H} D={D} seed={seed}", flush=True)
    rig = synth.rig_small(W, H, D, seed)
    ref = oracle_mod.compute(rig["views"], D, modes=7)
    gpu_ctx.vsweep(1)
    try:
        got = gpu_ctx.compute_rig(rig["views"], D, modes=7)
    finally:
        gpu_ctx.vsweep(-1)
    for k in ("multiview", "horizontal", "vertical", "disp"):
        assert np.array_equal(got[k], ref[k]), (k, int((got[k] != ref[k]).sum()))
    assert float(np.abs(got["subpix"] - ref["subpix"]).max()) <= 1e-3


# the fused sweep's partial sums kept over the rows family's bytes (the
# layout plans short of memory take): outputs, for rows whose last wave of
# columns is partial (W' not a multiple of the 64 / KL columns of a wave) and
# for several blocks, with 4 and 8 lanes per column
@pytest.mark.parametrize("W,H,D,seed,modes", [(96, 64, 64, 13, 7), (100, 44, 64, 17, 7), (68, 40, 128, 18, 7),
                                              (72, 40, 192, 19, 7), (400, 40, 64, 16, 7), (20, 16, 256, 38, 1),
                                              (12, 16, 384, 39, 1), (200, 16, 256, 37, 1)])
def test_fused_vertical_in_place_outputs(gpu_ctx, oracle_mod, W, H, D, seed, modes):
    print(f"[parity fused in place] W={W} H={H} D={D} seed={seed}", flush=True)
    rig = synth.rig_small(W, H, D, seed)
    ref = oracle_mod.compute(rig["views"], D, modes=modes)
    gpu_ctx.vsweep(1)
    gpu_ctx.sweep_in_place(True)
    try:
        got = gpu_ctx.compute_rig(rig["views"], D, modes=modes)
    finally:
        gpu_ctx.vsweep(-1)
        gpu_ctx.sweep_in_place(False)
    for k in ("multiview", "horizontal", "vertical", "disp"):
        if k in ref:
            assert np.array_equal(got[k], ref[k]), (k, int((got[k] != ref[k]).sum()))
    assert float(np.abs(got["subpix"] - ref["subpix"]).max()) <= 1e-3


# the fused sweep's 8-lane form (D in (192, 384]: 64 owned columns, 2 or 3
# chunks per lane): every stage (debug sums kept) and the outputs of its
# full-lane and partial forms, one and several blocks
@pytest.mark.parametrize("W,H,D,seed", [(16, 16, 200, 31), (16, 20, 256, 32), (16, 16, 264, 33)])
def test_stages_bit_exact_fused_vertical_8_lanes(gpu_ctx, oracle_mod, W, H, D, seed):
    print(f"[parity fused 8 lanes] W={W} H={H} D={D} seed={seed}", flush=True)
    gpu_ctx.vsweep(1)
    try:
        bad = _compare_stages(gpu_ctx, oracle_mod, W, H, D, seed)
    finally:
        gpu_ctx.vsweep(-1)
    assert not bad, bad


@pytest.mark.parametrize("W,H,D,seed", [(16, 16, 256, 34), (16, 16, 384, 35), (24, 16, 320, 36), (200, 16, 256, 37)])
def test_fused_vertical_8_lanes_outputs(gpu_ctx, oracle_mod, W, H, D, seed):
    print(f"[parity fused 8 lanes, outputs] W={W} H={H} D={D} seed={seed}", flush=True)
    rig = synth.rig_small(W, H, D, seed)
    ref = oracle_mod.compute(rig["views"], D, modes=1)
    gpu_ctx.vsweep(1)
    try:
        got = gpu_ctx.compute_rig(rig["views"], D, modes=1)
    finally:
        gpu_ctx.vsweep(-1)
    for k in ("multiview", "disp"):
        assert np.array_equal(got[k], ref[k]), (k, int((got[k] != ref[k]).sum()))
    assert float(np.abs(got["subpix"] - ref["subpix"]).max()) <= 1e-3


def test_plain_640x480_d64(gpu_ctx, oracle_mod):
    print("[parity] 640x480 D=64 seed=1", flush=True)
    rig = synth.rig_plain_640()
    ref = oracle_mod.compute(rig["views"], rig["D"], modes=7)
    got = gpu_ctx.compute_rig(rig["views"], rig["D"], modes=7)
    for k in ("multiview", "horizontal", "vertical", "disp"):
        assert np.array_equal(got[k], ref[k]), (k, int((got[k] != ref[k]).sum()))
    assert float(np.abs(got["subpix"] - ref["subpix"]).max()) <= 1e-3


# frames wider than 4096 pixels take the serial median's prepass form (8 slots
# per thread up to 8192, 16 beyond); the others sort the six fixed values in
# the serial kernel.  W' = 4136 and 8232 wide frames 0/180, H' = 4136 tall
# (so 4136-wide frames 90/270).
@pytest.mark.parametrize("W,H,D,seed", [(4104, 16, 16, 21), (16, 4104, 16, 22), (8200, 12, 8, 23)])
def test_stages_bit_exact_wide_frames(gpu_ctx, oracle_mod, W, H, D, seed):
    print(f"[parity wide] W={W} H={H} D={D} seed={seed}", flush=True)
    bad = _compare_stages(gpu_ctx, oracle_mod, W, H, D, seed)
    assert not bad, bad


# disparity counts beyond 192: wider SGM lane states, the tiled cost sum at 4
# groups of 64 (D = 200, 256) and the two-launch cost sum beyond, the row WTA's
# start-of-row segment when D > 256 (lane 63's leaving pixel before column 0),
# and D = 512, the largest accepted.  Tiny visible sizes keep the oracle's
# volumes in seconds (W' x H' x D up to 543 M cells at D = 512).
@pytest.mark.parametrize("W,H,D,seed", [(16, 16, 200, 81), (16, 16, 256, 82), (16, 16, 264, 83),
                                        (16, 12, 320, 84), (8, 8, 384, 85), (8, 4, 512, 86)])
def test_large_disparity_counts_outputs(gpu_ctx, oracle_mod, W, H, D, seed):
    print(f"[parity large D] W={W} H={H} D={D} seed={seed}", flush=True)
    rig = synth.rig_small(W, H, D, seed)
    ref = oracle_mod.compute(rig["views"], D, modes=1)
    got = gpu_ctx.compute_rig(rig["views"], D, modes=1)
    for k in ("multiview", "disp"):
        assert np.array_equal(got[k], ref[k]), (k, int((got[k] != ref[k]).sum()))
    assert float(np.abs(got["subpix"] - ref["subpix"]).max()) <= 1e-3


def test_stages_bit_exact_d264(gpu_ctx, oracle_mod):
    print("[parity stages] W=16 H=16 D=264 seed=87", flush=True)
    bad = _compare_stages(gpu_ctx, oracle_mod, 16, 16, 264, 87)
    assert not bad, bad


# the smallest rigs accepted (W, H multiples of 4, D = 8): 20 x 20 padded frames,
# smaller than one cost-sum tile, one mask tile or one WTA segment
@pytest.mark.parametrize("W,H,D,seed", [(4, 4, 8, 88), (8, 4, 8, 89)])
def test_smallest_rigs_bit_exact(gpu_ctx, oracle_mod, W, H, D, seed):
    print(f"[parity smallest] W={W} H={H} D={D} seed={seed}", flush=True)
    bad = _compare_stages(gpu_ctx, oracle_mod, W, H, D, seed)
    assert not bad, bad


def test_row_stride_host_and_device(gpu_ctx, oracle_mod):
    """Views inside wider rows (stride > W), as the reference's Mats of a
    larger image would be: the host form copies the rows out, the device
    form reads them in place.  The padding bytes differ from the image."""
    import ctypes as C

    import torch

    import sister_amd
    W, H, D, pitch = 64, 48, 16, 84
    rig = synth.rig_small(W, H, D, 90)
    ref = oracle_mod.compute(rig["views"], D, modes=1)
    print(f"[parity stride] W={W} H={H} D={D} stride={pitch}", flush=True)
    bufs = []
    for v in rig["views"]:
        b = np.full((H, pitch), 201, np.uint8)
        b[:, :W] = v
        bufs.append(b)
    r = sister_amd.RigStruct()
    r.width, r.height, r.stride, r.disp_count = W, H, pitch, D
    for i, b in enumerate(bufs):
        r.view[i] = b.ctypes.data
    mv = np.zeros((H, W), np.uint16)
    dp = np.zeros((H, W), np.uint16)
    r.out_multiview, r.out_disp = mv.ctypes.data, dp.ctypes.data
    gpu_ctx._check(sister_amd.lib().sister_compute_rig(gpu_ctx._h, C.byref(r), sister_amd.MODE_MULTIVIEW))
    assert np.array_equal(mv, ref["multiview"]) and np.array_equal(dp, ref["disp"])
    dev_bufs = [torch.from_numpy(b).cuda() for b in bufs]
    views = [t[:, :W] for t in dev_bufs]
    assert views[0].stride(0) == pitch
    outs = {"multiview": torch.zeros((H, W), dtype=torch.uint16, device="cuda"),
            "disp": torch.zeros((H, W), dtype=torch.uint16, device="cuda")}
    gpu_ctx.run_device([gpu_ctx.rig_struct(views, D, outs)], sister_amd.MODE_MULTIVIEW)
    gpu_ctx.sync()
    assert np.array_equal(outs["multiview"].cpu().numpy(), ref["multiview"])
    assert np.array_equal(outs["disp"].cpu().numpy(), ref["disp"])


# subsets of the three modes (hpp:82,87: horizontal uses frames 0/180 only,
# vertical 90/270 only); each output alone equals the oracle's
@pytest.mark.parametrize("modes", [2, 4, 6, 3, 5])
def test_mode_subsets(gpu_ctx, oracle_mod, modes):
    W, H, D, seed = 96, 64, 24, 92
    print(f"[parity modes] modes={modes} W={W} H={H} D={D}", flush=True)
    rig = synth.rig_small(W, H, D, seed)
    ref = oracle_mod.compute(rig["views"], D, modes=modes)
    got = gpu_ctx.compute_rig(rig["views"], D, modes=modes)
    names = [n for b, n in ((1, "multiview"), (2, "horizontal"), (4, "vertical")) if modes & b]
    assert sorted(k for k in got if k in ("multiview", "horizontal", "vertical")) == sorted(names)
    for k in names + (["disp"] if modes & 1 else []):
        assert np.array_equal(got[k], ref[k]), (k, int((got[k] != ref[k]).sum()))
