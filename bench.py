#!/usr/bin/env python3
"""Benchmark of the 5-view disparity path on one MI355X.

Default workload: 1920x1080 rigs, D=192 (the reference README's setting),
mode 0 (multiview).  Two forms are measured on the same synthetic rigs, with
inputs and outputs resident in HBM before the timed region:

  plain    one rig per call of the library (sister_run_device with n = 1),
  batched  B rigs in one plan (default 48: 182 GB of workspace), captured
           once as a hipGraph and replayed.

`value` is the batched throughput in G disparity-evals/s (sum of W*H*D over
the rigs of a step / step time); the plain figures and the ratio sit beside
it.  Per-kernel device times come from HIP events on the library stream
(instrumented eager steps of the same batch before the graph is captured).
The roofline entry prices the dominant kernel's algorithmic bytes against the
MI355X HBM peak (and, in roofline.valu, its integer work against the VALU
peak).  The CPU baseline times the reference's own kernels
(oracle/_ref, census.cpp / sgm.cpp / postprocess.cpp) on one rig.

    python bench.py [--steps K] [--warmup W] [--batch B] [--workload rig1080|mixed32|large|serve]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("G disparity-evals/s and rigs/s (p50/p99 latency), batched vs plain, "
          "1 MI355X, bit-exact")
VALU_PEAK = 256 * 64 * 2.4e9      # int32 lane-ops/s: 256 CUs x 64 lanes/clk x 2.4 GHz
HBM_PEAK = 8.0e12

# Algorithmic integer lane-ops per padded cell (W'*H'*D) of mode 0, split per
# kernel: 4 directional costs (xor 2 + popcount 2 + invalid select 1 each = 20)
# and the 8 LR-check argmins live in `wta`; the mask-weighted sum is 4; each
# SGM path is 8 (+P1, min d-1, min d+1, min P2 term, -minL, +C, running min,
# +S) and each chain-family kernel sweeps two paths; the final argmin (1) is
# fused into sgm_anti.  Total 28 + 4 + 3*16 + 17 = 97.
OPS_PER_CELL = {"wta": 28, "msum": 4, "sgm_rows": 16, "sgm_cols": 16, "sgm_diag": 16, "sgm_anti": 17,
                "sgm_vsweep": 49}   # the fused vertical sweep: 3 families x 2 sweeps x 8, + argmin

# Algorithmic HBM bytes per padded cell of the HBM-bound kernels.  Each SGM
# chain family reads the u8 cost M in both sweeps (2), writes its forward
# (L - M) byte (1), reads it back and writes the two-direction sum (2): 5.
# The anti-diagonal family, which forms S and the winner, instead reads the
# three other families' bytes (3) and writes nothing in its backward sweep: 7.
# Ahead of the fused vertical sweep the rows family stores its two directions
# separately (read M twice, write two bytes): 4.  The fused vertical sweep
# (columns, diagonals, anti-diagonals together) reads M (1) and the rows
# family's two bytes (2) and writes the u16 partial sum P (2) going down,
# then reads M (1) and P (2) going up: 8 (against 5 + 5 + 7 for the three
# separate families).  The cost sum over 16x16 tiles writes M once (1); its
# two-launch form (frames 0/180, then 90/270 read-modify-write) moves 3.
BYTES_PER_CELL = {"sgm_rows": 4, "sgm_cols": 5, "sgm_diag": 5, "sgm_anti": 7, "sgm_vsweep": 8, "msum": 1}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_batch48_rig1080_d192.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def padded_cells(W, H, D):
    return (W + 2 * D) * (H + 2 * D) * D


def make_rigs(workload, batch):
    from sister_amd import synth
    if workload == "rig1080":
        scenes = [synth.rig_scene(1920, 1080, 192, 2 + i, n_boxes=5, d_plane=(20.0, 60.0), d_box=(60.0, 180.0))
                  for i in range(min(batch, 4))]
        return [scenes[i % len(scenes)] for i in range(batch)]
    if workload == "mixed32":
        return synth.rig_batch(32, 1000)
    if workload == "large":
        return [synth.rig_large(5)]
    raise ValueError(workload)


def to_device(ctx, rigs, torch):
    dev = []
    for r in rigs:
        views = [torch.from_numpy(v).cuda() for v in r["views"]]
        H, W = r["views"][0].shape
        outs = {"multiview": torch.zeros((H, W), dtype=torch.uint16, device="cuda"),
                "disp": torch.zeros((H, W), dtype=torch.uint16, device="cuda"),
                "subpix": torch.zeros((H, W), dtype=torch.float32, device="cuda")}
        dev.append(dict(views=views, outs=outs, D=r["D"], struct=ctx.rig_struct(views, r["D"], outs)))
    return dev


def cpu_baseline(rig):
    """Reference kernels (oracle/_ref) on one rig, mode 0; the scalar port
    when the reference build is absent."""
    from oracle import oracle as O
    kind = "reference" if O.ref_available() else "port"
    t = time.perf_counter()
    O.compute(rig["views"], rig["D"], modes=1, use_ref=(kind == "reference"), subpix=False)
    dt = time.perf_counter() - t
    H, W = rig["views"][0].shape
    return {
        "value": W * H * rig["D"] / dt / 1e9,
        "unit": "G disparity-evals/s",
        "cores": 4 if kind == "reference" else 1,
        "kind": kind,
        "sample": (f"one {W}x{H} D={rig['D']} rig, doMultiStereo mode 0 (4 cost volumes, masks, "
                   f"weighted sum, SGM, WTA) in {dt:.2f} s wall; the reference's OpenMP uses 4 threads "
                   f"in the cost stage only"),
        "seconds": dt,
    }


def run_single(args):
    """One instrumented eager run of the batch, the same launches whose event
    times give kernel_ms (one dispatch per kernel and D group, families in
    sequence): the command profiled by rocprofv3 (--kernel-trace, --pmc)."""
    import torch
    import sister_amd

    ctx = sister_amd.Context(device=0)
    rigs = make_rigs(args.workload, args.batch)
    dev = to_device(ctx, rigs, torch)
    torch.cuda.synchronize()
    ctx.profile(True)
    ctx.run_device([d["struct"] for d in dev], 1)
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    cells = sum(padded_cells(r["W"], r["H"], r["D"]) for r in rigs)
    print(json.dumps({"single_run_rigs": len(rigs), "padded_cells": cells,
                      "kernel_ms": {k: ms for k, (ms, n) in prof.items() if n}}))
    ctx.close()


def run(args):
    import numpy as np
    import torch
    import sister_amd

    ctx = sister_amd.Context(device=0)
    rigs = make_rigs(args.workload, args.batch)
    dev = to_device(ctx, rigs, torch)
    B = len(dev)
    evals = sum(r["views"][0].shape[0] * r["views"][0].shape[1] * r["D"] for r in rigs)
    cells = sum(padded_cells(r["W"], r["H"], r["D"]) for r in rigs)
    torch.cuda.synchronize()

    # ---- plain: one rig per call ------------------------------------------
    for i in range(args.warmup):
        ctx.run_device([dev[i % B]["struct"]], 1)
    ctx.sync()
    torch.cuda.synchronize()
    nplain = max(args.steps, 1)
    t0 = time.perf_counter()
    for i in range(nplain):
        ctx.run_device([dev[i % B]["struct"]], 1)
    ctx.sync()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / nplain
    plain_evals = sum(rigs[i % B]["W"] * rigs[i % B]["H"] * rigs[i % B]["D"] for i in range(nplain)) / nplain
    ncmp = min(B, 4)
    for i in range(nplain, ncmp):          # every compared rig computed alone once
        ctx.run_device([dev[i]["struct"]], 1)
    ctx.sync()
    plain_out = [d["outs"]["disp"].cpu().numpy().copy() for d in dev[:ncmp]]

    # ---- per-kernel device time: instrumented eager steps of the same batch --
    # (before the capture; a one-rig call then replaces the batch-sized eager
    # plan, so that only the graph's workspaces are held while it is timed)
    # (cost-sum and SGM launches run once per group of equal D: the figures
    # are per-step totals over a kernel's launches)
    nrep = max(1, min(args.steps, 3))
    ctx.profile(True)
    for _ in range(nrep):
        ctx.run_device([d["struct"] for d in dev], 1)
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    ctx.run_device([dev[0]["struct"]], 1)
    ctx.sync()

    # ---- batched: all B rigs in one captured plan ----------------------------
    for d in dev:
        d["outs"]["disp"].zero_()
    ctx.graph_capture([d["struct"] for d in dev], 1)
    for _ in range(args.warmup):
        ctx.graph_launch()
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.graph_launch()
    ctx.sync()
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    same = all(np.array_equal(dev[i]["outs"]["disp"].cpu().numpy(), plain_out[i]) for i in range(len(plain_out)))

    kern = {k: ms / nrep for k, (ms, n) in prof.items() if n}
    launches = {k: n / nrep for k, (ms, n) in prof.items() if n}
    dom = max(kern, key=lambda k: kern[k])
    if dom in BYTES_PER_CELL:
        achieved = BYTES_PER_CELL[dom] * cells / (kern[dom] * 1e-3)
        traffic = None
        if os.path.exists(TRAFFIC_FILE):
            t = json.load(open(TRAFFIC_FILE))
            matches = (args.workload == "rig1080" and B == t.get("rigs"))
            k = t["kernels"].get(dom)
            if matches and k:
                traffic = k["fetch_bytes"] + k["write_bytes"]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                    "bytes_per_cell": BYTES_PER_CELL[dom], "launches_per_step": launches[dom]}
        if dom in OPS_PER_CELL:
            # the same kernel priced by its algorithmic integer work
            v = OPS_PER_CELL[dom] * cells / (kern[dom] * 1e-3)
            roofline["valu"] = {"achieved": v / 1e12, "peak": VALU_PEAK / 1e12, "unit": "T int-lane-ops/s",
                                "frac": v / VALU_PEAK, "ops_per_cell": OPS_PER_CELL[dom]}
    else:
        achieved = OPS_PER_CELL.get(dom, 0) * cells / (kern[dom] * 1e-3)
        roofline = {"bound": "valu", "kernel": dom, "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12,
                    "unit": "T int-lane-ops/s", "frac": achieved / VALU_PEAK, "traffic": None,
                    "ops_per_cell": OPS_PER_CELL.get(dom, 0)}

    out = {
        "metric": METRIC,
        "value": evals / (step_ms * 1e-3) / 1e9,
        "unit": "G disparity-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic (seeded noise scenes, plane + boxes, backward-warped neighbours)",
        "config": {
            "workload": (f"{args.workload}: {B} rigs per step in one hipGraph-captured plan"
                         if args.workload != "rig1080" else
                         f"1920x1080 D=192 5-view rigs, multiview mode; batched = {B} rigs per hipGraph "
                         f"launch, plain = 1 rig per call"),
            "rigs_per_step": B,
            "modes": "multiview",
        },
        "rigs_per_s": B / (step_ms * 1e-3),
        "plain": {"ms_per_rig": plain_ms, "rigs_per_s": 1e3 / plain_ms,
                  "value": plain_evals / (plain_ms * 1e-3) / 1e9},
        "batched_vs_plain": (B / step_ms) / (1.0 / plain_ms),
        "batched_equals_plain": bool(same),
        "padded_cells_per_step": cells,
        "pipeline_valu_frac": 97 * cells / (step_ms * 1e-3) / VALU_PEAK,
        "roofline": roofline,
        "kernel_ms": kern,
        "kernel_launches_per_step": launches,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(rigs[0])
    ctx.close()
    return out


def run_serve(args):
    """Continuous batching under Poisson arrivals: the mixed-size pool (640x480
    to 2048x1536, D 64..256) plus 5% 4096x3000 rigs, HBM budget --hbm-gb.
    Default: device-resident buffers (sister_submit_device); --serve-host
    submits host buffers (staged in and out by the server)."""
    import numpy as np
    import torch
    import sister_amd
    from sister_amd import synth

    ctx = sister_amd.Context(device=0, hbm_budget=args.hbm_gb << 30)
    shapes = synth.batch_shapes(64, 1000)
    pool = [synth.rig_scene(W, H, D, 1000 + i) for i, (W, H, D) in enumerate(shapes[:8])]
    large = synth.rig_scene(4096, 3000, shapes[0][2], 5, n_boxes=20, blur_sigma=0.0)
    rng = np.random.default_rng(7)
    n = int(args.rate * args.serve_seconds)
    # request i uses the large rig with probability 0.05, else a pool rig
    picks = [-1 if rng.random() < 0.05 else int(rng.integers(0, len(pool))) for _ in range(n)]
    gaps = rng.exponential(1.0 / args.rate, size=n)
    rig_of = lambda k: large if k < 0 else pool[k]
    device = not args.serve_host
    slots = {}   # (pick, slot) -> (struct, keep-alive); two output sets per rig
    for k in sorted(set(picks)):
        r = rig_of(k)
        for sl in range(2):
            if device:
                d = to_device(ctx, [r], torch)[0]
                slots[(k, sl)] = (d["struct"], d)
            else:
                o = np.zeros((r["H"], r["W"]), np.uint16)
                st = sister_amd.RigStruct()
                st.width, st.height, st.stride, st.disp_count = r["W"], r["H"], r["W"], r["D"]
                for v in range(5):
                    st.view[v] = r["views"][v].ctypes.data
                st.out_multiview = o.ctypes.data
                slots[(k, sl)] = (st, o)
    torch.cuda.synchronize()
    ctx.serve_start(4)
    tickets = []
    t0 = time.perf_counter()
    tnext = t0
    tlog = t0
    for i, k in enumerate(picks):
        while time.perf_counter() < tnext:
            time.sleep(0.0002)
        tickets.append(ctx.submit(slots[(k, i & 1)][0], 1, device=device))
        tnext += gaps[i]
        if time.perf_counter() - tlog > 2.0:
            tlog = time.perf_counter()
            log(f"serve: {i + 1}/{n} submitted, {ctx.stats()['completed']} completed")
    failed, reason = 0, None
    for i, t in enumerate(tickets):
        try:
            ctx.wait(t)
        except sister_amd.SisterError as e:
            failed += 1
            reason = str(e)
        if time.perf_counter() - tlog > 2.0:
            tlog = time.perf_counter()
            log(f"serve: {i + 1}/{n} completed")
    wall = time.perf_counter() - t0
    st = ctx.stats()
    ctx.serve_stop()
    ctx.close()
    evals = sum(rig_of(k)["W"] * rig_of(k)["H"] * rig_of(k)["D"] for k in picks)
    return {
        "metric": METRIC, "value": evals / wall / 1e9, "unit": "G disparity-evals/s", "n_gpus": 1,
        "steps": n, "warmup": 0, "ms_per_step": wall * 1e3 / n, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": f"serving: Poisson {args.rate} rigs/s for {args.serve_seconds} s, mixed sizes/D "
                               f"+ 5% 4096x3000, {'device-resident' if device else 'host'} buffers, "
                               f"HBM budget {args.hbm_gb} GiB"},
        "rigs_per_s": n / wall, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
        "large_rigs": sum(1 for k in picks if k < 0), "failed": failed, "fail_reason": reason, "stats": st,
    }


def calibrate_copy(gib=1):
    """A plain streaming kernel of known traffic (reads and writes `gib` GiB)
    for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = (gib << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        torch.add(a, 1.0, out=b)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 3
    print(json.dumps({"calibrate_copy_bytes_read": n * 4, "calibrate_copy_bytes_written": n * 4, "launches": 3,
                      "ms_per_launch": ms, "copy_GBps": 2 * n * 4 / (ms * 1e-3) / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=48)
    ap.add_argument("--workload", default="rig1080", choices=["rig1080", "mixed32", "large", "serve"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-host", action="store_true", help="serve host buffers instead of device-resident ones")
    ap.add_argument("--hbm-gb", type=int, default=64)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--single", action="store_true", help="one eager run of the batch (counter runs)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.single:
        run_single(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only; --gpus must be 1 (rigs are independent, "
                                   "a multi-GPU run would shard rigs across processes)"}))
        sys.exit(2)
    out = run_serve(args) if args.workload == "serve" else run(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
