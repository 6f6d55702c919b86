#!/usr/bin/env python3
"""Benchmark of the 5-view disparity path on one MI355X.

Default workload: 1920x1080 rigs, D=192 (the reference README's setting),
mode 0 (multiview).  Two forms are measured on the same synthetic rigs, with
inputs and outputs resident in HBM before the timed region:

  plain    one rig per call of the library (sister_run_device with n = 1),
  batched  B rigs in one plan (default 64: 229 GB of workspace), captured
           once as a hipGraph and replayed.

`value` is the batched throughput in G disparity-evals/s (sum of W*H*D over
the rigs of a step / step time); the plain figures and the ratio sit beside
it, and every rig of the batch is compared with its plain result.

`roofline` prices the whole step by its algorithmic integer work: 97 int
lane-ops per padded cell (W'*H'*D) against the MI355X VALU peak, with the HBM
bytes the counters measured per step beside it.  Its `dominant_kernel` entry
does the same for the kernel with the most device time: its own
ops per cell, its average launch duration (HIP events around its launches on
the library stream, in instrumented eager steps of the same batch taken
before the graph is captured; real dispatch counts), and its counter traffic
against its algorithmic bytes.

Two more blocks run after the headline, each in a fresh context:
  mixed32  32 rigs of mixed sizes (640x480 .. 2048x1536) and D (64 .. 256)
           captured as one graph;
  serve    Poisson arrivals at 200 rigs/s from the same mix plus 5%
           4096x3000 rigs, HBM budget 64 GiB: rigs/s, p50/p99 latency.

The CPU baseline times the reference's own kernels (oracle/_ref:
census.cpp / sgm.cpp / postprocess.cpp behind the restated glue) on the
host: with the reference's parallelism (OpenMP 4 threads in the cost stage),
one warm-up and the median of three runs of one 1920x1080 D=192 rig; and on
every core, one rig-sized band per core computed concurrently.

    python bench.py [--steps K] [--warmup W] [--batch B]
                    [--workload rig1080|mixed32|large|serve] [--no-extra]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("G disparity-evals/s and rigs/s (p50/p99 latency), batched vs plain, "
          "1 MI355X, bit-exact")
VALU_PEAK = 256 * 64 * 2.4e9      # int32 lane-ops/s: 256 CUs x 64 lanes/clk x 2.4 GHz
HBM_PEAK = 8.0e12
OPS_PER_CELL_PIPELINE = 97        # mode 0: costs 20, LR-check argmins 8, weighted sum 4, SGM 64, WTA 1

# Algorithmic integer lane-ops per padded cell (W'*H'*D) of mode 0, split per
# kernel: 4 directional costs (xor 2 + popcount 2 + invalid select 1 each = 20)
# and the 8 LR-check argmins live in `wta`; the mask-weighted sum is 4; each
# SGM path is 8 (+P1, min d-1, min d+1, min P2 term, -minL, +C, running min,
# +S) and each chain-family kernel sweeps two paths; the final argmin (1) is
# fused into sgm_anti.  Total 28 + 4 + 3*16 + 17 = 97.
OPS_PER_CELL = {"wta": 28, "msum": 4, "sgm_rows": 16, "sgm_cols": 16, "sgm_diag": 16, "sgm_anti": 17,
                "sgm_vsweep": 49}   # the fused vertical sweep: 3 families x 2 sweeps x 8, + argmin

# Algorithmic HBM bytes per padded cell of the HBM-bound kernels, with
# `crop` the fraction of padded cells inside the crop (W*H of W'*H'): the
# intermediates whose only reader is the final aggregate S are moved for those
# cells alone, since S is read at the crop only.  Each separate SGM chain
# family reads the u8 cost M in both sweeps (2), writes its forward (L - M)
# byte (1), reads it back and writes the two-direction sum (2): 5.  The
# anti-diagonal family, which forms S and the winner, instead reads the three
# other families' bytes (3) and writes nothing in its backward sweep: 7.  The
# rows family sweeps its two directions as separate waves, each reading M (2),
# and stores one byte per direction at crop cells (2 crop).  The fused
# vertical sweep (columns, diagonals, anti-diagonals together) reads M going
# down and again going up (2) and, at crop cells, the rows family's two bytes
# (2), the u16 partial sum P written going down (2) and read going up (2):
# 2 + 6 crop (against 5 + 5 + 7 for the three separate families).  The cost
# sum over 16x16 tiles writes M once (1); its two-launch form (frames 0/180,
# then 90/270 read-modify-write) moves 3.
def bytes_per_cell(kernel, crop):
    fixed = {"sgm_cols": 5, "sgm_diag": 5, "sgm_anti": 7, "msum": 1}
    if kernel == "sgm_rows":
        return 2 + 2 * crop
    if kernel == "sgm_vsweep":
        return 2 + 6 * crop
    return fixed.get(kernel)


# per-kernel counter totals of one instrumented step (used when their batch
# size is the one measured)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_rig1080_d192.json")
VALU_FILE = os.path.join(ROOT, "profiles", "valu_rig1080_d192.json")


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


# ---------------------------------------------------------------------------
# CPU baseline
# ---------------------------------------------------------------------------
def host_description():
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable"):
                    avail = int(line.split()[1]) * 1024
    except OSError:
        pass
    return {"cpus_usable": cores, "cpus_online": os.cpu_count(), "model": model, "mem_available": avail}


def cpu_baseline(rig):
    """The reference's kernels (oracle/_ref) behind the restated glue, mode 0,
    timed on the host.  Reference parallelism: OpenMP 4 threads in the cost
    stage only (census.cpp:116-139, numThreads=4 at hpp:77), every other
    stage single-threaded; one warm-up on a 640x480 D=64 rig, then the median
    of three runs of the 1920x1080 D=192 rig.  All cores: one band of the
    same rig (all 1920 columns, 128 rows, D=192) per usable core, computed
    concurrently (a rig per core, the batch's own parallelism), scaled to
    the full rig's evals per padded cell.  The scalar port stands in when the
    reference build is absent."""
    from oracle import oracle as O
    from sister_amd import synth
    kind = "reference" if O.ref_available() else "port"
    use_ref = kind == "reference"
    host = host_description()
    H, W, D = rig["H"], rig["W"], rig["D"]
    warm = synth.rig_small(640, 480, 64, 1)
    t = time.perf_counter()
    O.compute(warm["views"], 64, modes=1, use_ref=use_ref, subpix=False)
    warm_s = time.perf_counter() - t
    times = []
    for _ in range(3):
        t = time.perf_counter()
        O.compute(rig["views"], D, modes=1, use_ref=use_ref, subpix=False)
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    evals = W * H * D
    evals_per_cell = evals / padded_cells(W, H, D)
    out = {
        "value": evals / med / 1e9,
        "unit": "G disparity-evals/s",
        "cores": 4 if use_ref else 1,
        "kind": kind,
        "sample": (f"one {W}x{H} D={D} rig, doMultiStereo mode 0 (4 cost volumes, masks, weighted sum, SGM, "
                   f"WTA), median of 3 runs after a 640x480 D=64 warm-up; the reference's OpenMP uses 4 threads "
                   f"in the cost stage only, the other stages one"),
        "seconds": times,
        "warmup_seconds": warm_s,
        "host": host,
    }
    # all cores: bands of the same scene, one per usable core, bounded by memory
    bh = 128
    per_copy = padded_cells(W, bh, D) * 2 * 9 + 16 * W * bh   # nine u16 volumes plus images
    ncores = host["cpus_usable"] or 1
    ncopies = min(ncores, 16)   # at most the 16 cores a one-GPU share of the host holds
    if host["mem_available"]:
        ncopies = max(1, min(ncores, int(host["mem_available"] * 0.4 // per_copy)))
    band = [v[:bh] for v in rig["views"]]
    t_each = [0.0] * ncopies

    def work(k):
        t0 = time.perf_counter()
        O.compute(band, D, modes=1, use_ref=use_ref, subpix=False)
        t_each[k] = time.perf_counter() - t0

    th = [threading.Thread(target=work, args=(k,)) for k in range(ncopies)]
    t0 = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    wall = time.perf_counter() - t0
    cells_rate = ncopies * padded_cells(W, bh, D) / wall
    out["all_cores"] = {
        "value": cells_rate * evals_per_cell / 1e9,
        "unit": "G disparity-evals/s (at the 1920x1080 D=192 rig's evals per padded cell)",
        "padded_cells_per_s": cells_rate,
        "threads": ncopies,
        "sample": (f"{ncopies} concurrent {W}x{bh} D={D} bands of the same scene, each with the reference's "
                   f"parallelism ({ncores} usable cores; copies limited to 16, the one-GPU share of a host, and "
                   f"to 40% of the available memory), one round after the runs above, {wall:.2f} s wall"),
        "seconds_per_copy": [round(x, 3) for x in t_each],
    }
    return out


# ---------------------------------------------------------------------------
# profiling run (the command rocprofv3 wraps)
# ---------------------------------------------------------------------------
def run_single(args):
    """One instrumented eager run of the batch, the same launches whose event
    times give kernel_ms (families in sequence): the command profiled by
    rocprofv3 (--kernel-trace, --pmc)."""
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
                      "kernel_ms": {k: ms for k, (ms, n) in prof.items() if n},
                      "kernel_dispatches": {k: n for k, (ms, n) in prof.items() if n}}))
    ctx.close()


# ---------------------------------------------------------------------------
# headline: plain vs batched at 1920x1080 D=192
# ---------------------------------------------------------------------------
def wta_copied_fraction(rigs):
    """Fraction of the WTA's cost evaluations (padded cells x 4 frames) that
    fall on the deep padding bands' rows, [3, D-4] and [R-D+3, R-4] of every
    frame, which copy the band's first row instead of computing it."""
    done = total = 0
    for r in rigs:
        W, H, D = r["W"], r["H"], r["D"]
        Wp, Hp = W + 2 * D, H + 2 * D
        # copied: 2 bands x (D-7) rows x (2 W' + 2 H') columns x D, of 4 W' H' D
        total += Wp * Hp * D
        done += (D - 7) * (Wp + Hp) * D
    return done / total


def roofline_entry(kern, launches, cells, step_ms, B, workload, crop, wta_copied=0.0):
    traffic = None
    per_kernel_traffic = {}
    valu = {}
    if os.path.exists(TRAFFIC_FILE):
        t = json.load(open(TRAFFIC_FILE))
        if workload == "rig1080" and B == t.get("rigs"):
            per_kernel_traffic = {k: v["fetch_bytes"] + v["write_bytes"] for k, v in t["kernels"].items()}
            traffic = sum(per_kernel_traffic.values())
    if os.path.exists(VALU_FILE):
        v = json.load(open(VALU_FILE))
        if workload == "rig1080" and B == v.get("rigs"):
            valu = v.get("kernels", {})
    step_s = step_ms * 1e-3
    achieved = OPS_PER_CELL_PIPELINE * cells / step_s
    # the same without the WTA evaluations of the deep padding bands, which
    # the kernels copy rather than compute
    executed = OPS_PER_CELL_PIPELINE - OPS_PER_CELL["wta"] * wta_copied
    roof = {"bound": "valu", "scope": "whole step (every kernel of the batch)",
            "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12, "unit": "T int-lane-ops/s",
            "frac": achieved / VALU_PEAK, "ops_per_cell": OPS_PER_CELL_PIPELINE,
            "ops_per_cell_computed": executed,
            "frac_computed": executed * cells / step_s / VALU_PEAK,
            "traffic": traffic,
            "hbm_achieved": traffic / step_s / 1e9 if traffic else None,
            "hbm_frac": traffic / step_s / HBM_PEAK if traffic else None,
            "hbm_unit": "GB/s"}
    dom = max(kern, key=lambda k: kern[k])
    n = launches.get(dom) or 1
    avg_ms = kern[dom] / n
    ops = OPS_PER_CELL.get(dom, 0) * cells
    d = {"kernel": dom, "bound": "valu", "ops_per_cell": OPS_PER_CELL.get(dom, 0),
         "launches_per_step": n, "avg_launch_ms": avg_ms,
         "achieved": ops / n / (avg_ms * 1e-3) / 1e12, "peak": VALU_PEAK / 1e12, "unit": "T int-lane-ops/s",
         "frac": ops / n / (avg_ms * 1e-3) / VALU_PEAK}
    bpc = bytes_per_cell(dom, crop)
    if bpc is not None:
        alg = bpc * cells
        d.update({"bytes_per_cell": bpc, "crop_fraction": crop,
                  "hbm_achieved": alg / (kern[dom] * 1e-3) / 1e9,
                  "hbm_frac": alg / (kern[dom] * 1e-3) / HBM_PEAK, "hbm_unit": "GB/s (algorithmic bytes)"})
        if dom in per_kernel_traffic:
            d["traffic"] = per_kernel_traffic[dom]
            d["traffic_over_algorithmic"] = per_kernel_traffic[dom] / alg
    if dom in valu:
        d["valu_counters"] = valu[dom]
    roof["dominant_kernel"] = d
    return roof


def time_graph(ctx, args, torch):
    """Replays of the context's captured graph: warm-up, then the timed steps."""
    for _ in range(args.warmup):
        ctx.graph_launch()
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.graph_launch()
    ctx.sync()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / args.steps


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
    log(f"headline: {B} rigs, {cells / 1e9:.2f} G padded cells per step")

    # warm-up of the one-rig path
    for i in range(args.warmup):
        ctx.run_device([dev[i % B]["struct"]], 1)
    ctx.sync()

    # ---- per-kernel device time: instrumented eager steps of the same batch --
    # (before the capture; a one-rig call then replaces the batch-sized eager
    # plan, so that only the graph's workspaces are held while it is timed)
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
    # first in the default launch forms, then in the forms the autotuner picks
    # for this batch (graph parts; per disparity group the SGM and cost-sum
    # forms), which is what is reported
    for d in dev:
        d["outs"]["disp"].zero_()
        d["outs"]["multiview"].zero_()
    structs = [d["struct"] for d in dev]
    ctx.graph_capture(structs, 1)
    default_ms = time_graph(ctx, args, torch)
    log(f"batched, default forms: {default_ms:.2f} ms per step")
    t0 = time.perf_counter()
    ctx.autotune(structs, 1)
    tune_s = time.perf_counter() - t0
    for d in dev:
        d["outs"]["disp"].zero_()
        d["outs"]["multiview"].zero_()
    step_ms = time_graph(ctx, args, torch)
    log(f"batched, tuned forms: {step_ms:.2f} ms per step (tuning {tune_s:.1f} s)")
    batched_out = [{k: d["outs"][k].clone() for k in ("disp", "multiview")} for d in dev]

    # ---- plain: one rig per call, timed right after the batched steps -------
    # (under the same conditions as the reported batched figure)
    nplain = max(args.steps, 1)
    ctx.run_device([dev[0]["struct"]], 1)
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(nplain):
        ctx.run_device([dev[i % B]["struct"]], 1)
    ctx.sync()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / nplain
    plain_evals = sum(rigs[i % B]["W"] * rigs[i % B]["H"] * rigs[i % B]["D"] for i in range(nplain)) / nplain
    log(f"plain: {plain_ms:.2f} ms per rig")
    # the distinct scenes (rig i is scene i % nscene), each computed alone
    nscene = min(B, 4) if args.workload == "rig1080" else B
    for i in range(nscene):
        ctx.run_device([dev[i]["struct"]], 1)
        ctx.sync()
    plain_out = [{k: dev[i]["outs"][k].clone() for k in ("disp", "multiview")} for i in range(nscene)]
    bad = [i for i in range(B)
           if not all(torch.equal(batched_out[i][k], plain_out[i % nscene][k]) for k in ("disp", "multiview"))]

    kern = {k: ms / nrep for k, (ms, n) in prof.items() if n}
    launches = {k: n / nrep for k, (ms, n) in prof.items() if n}
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
        "default_forms_ms_per_step": default_ms,
        "autotune": {"seconds": tune_s, "forms": ctx.tune_table().strip().splitlines()[1:],
                     "note": "kernel_ms comes from eager steps in the default forms"},
        "batched_equals_plain": not bad,
        "batched_mismatched_rigs": len(bad),
        "batched_compared": f"disp and multiview maps of all {B} rigs against their scene computed alone",
        "padded_cells_per_step": cells,
        "padded_cells_per_s": cells / (step_ms * 1e-3),
        "roofline": roofline_entry(kern, launches, cells, step_ms, B, args.workload, evals / cells,
                                   wta_copied_fraction(rigs)),
        "kernel_ms": kern,
        "kernel_launches_per_step": launches,
    }
    ctx.close()
    del dev
    torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(rigs[0])
    return out


def run_mixed(args):
    """The mixed batch: 32 rigs (sizes 640x480 .. 2048x1536, D 64 .. 256) in
    one captured graph."""
    import torch
    import sister_amd

    ctx = sister_amd.Context(device=0)
    rigs = make_rigs("mixed32", 32)
    dev = to_device(ctx, rigs, torch)
    evals = sum(r["W"] * r["H"] * r["D"] for r in rigs)
    cells = sum(padded_cells(r["W"], r["H"], r["D"]) for r in rigs)
    torch.cuda.synchronize()
    structs = [d["struct"] for d in dev]
    # each rig alone, for the comparison below
    alone = []
    for d in dev:
        ctx.run_device([d["struct"]], 1)
        ctx.sync()
        alone.append(d["outs"]["disp"].clone())
    ctx.graph_capture(structs, 1)
    default_ms = time_graph(ctx, args, torch)
    log(f"mixed32, default forms: {default_ms:.2f} ms per step")
    t0 = time.perf_counter()
    ctx.autotune(structs, 1)
    tune_s = time.perf_counter() - t0
    for d in dev:
        d["outs"]["disp"].zero_()
    step_ms = time_graph(ctx, args, torch)
    bad = sum(1 for d, a in zip(dev, alone) if not torch.equal(d["outs"]["disp"], a))
    forms = ctx.tune_table().strip().splitlines()[1:]
    ctx.close()
    del dev
    torch.cuda.empty_cache()
    log(f"mixed32, tuned forms: {step_ms:.2f} ms per step (tuning {tune_s:.1f} s)")
    return {"workload": "32 rigs, sizes uniform over {640x480, 1280x720, 1920x1080, 2048x1536}, D uniform over "
                        "{64, 128, 192, 256} (seed 1000), one hipGraph",
            "ms_per_step": step_ms, "rigs_per_s": 32 / (step_ms * 1e-3),
            "value": evals / (step_ms * 1e-3) / 1e9, "unit": "G disparity-evals/s",
            "padded_cells_per_s": cells / (step_ms * 1e-3), "steps": args.steps,
            "default_forms_ms_per_step": default_ms, "autotune_seconds": tune_s, "forms": forms,
            "rigs_differing_from_alone": bad}


def run_serve(args, seconds=None):
    """Continuous batching under Poisson arrivals: the mixed-size pool (640x480
    to 2048x1536, D 64..256) plus 5% 4096x3000 rigs, HBM budget --hbm-gb.
    Default: device-resident buffers (sister_submit_device); --serve-host
    submits host buffers (staged in and out by the server)."""
    import numpy as np
    import torch
    import sister_amd
    from sister_amd import synth

    seconds = seconds or args.serve_seconds
    ctx = sister_amd.Context(device=0, hbm_budget=args.hbm_gb << 30)
    shapes = synth.batch_shapes(64, 1000)
    pool = [synth.rig_scene(W, H, D, 1000 + i) for i, (W, H, D) in enumerate(shapes[:8])]
    large = synth.rig_scene(4096, 3000, shapes[0][2], 5, n_boxes=20, blur_sigma=0.0)
    rng = np.random.default_rng(7)
    n = int(args.rate * seconds)
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
    wall = time.perf_counter() - t0
    st = ctx.stats()
    ctx.serve_stop()
    ctx.close()
    slots.clear()
    torch.cuda.empty_cache()
    evals = sum(rig_of(k)["W"] * rig_of(k)["H"] * rig_of(k)["D"] for k in picks)
    log(f"serve: {n / wall:.1f} rigs/s, p50 {st['p50_ms']:.1f} ms, p99 {st['p99_ms']:.1f} ms")
    return {
        "workload": (f"serving: Poisson {args.rate:g} rigs/s for {seconds:g} s, mixed sizes/D + 5% 4096x3000, "
                     f"{'device-resident' if device else 'host'} buffers, HBM budget {args.hbm_gb} GiB"),
        "requests": n, "rigs_per_s": n / wall, "value": evals / wall / 1e9, "unit": "G disparity-evals/s",
        "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
        "large_rigs": sum(1 for k in picks if k < 0), "failed": failed, "fail_reason": reason,
        "rejected": st["rejected"], "hbm_peak": st["hbm_peak"],
    }


def serve_line(args):
    s = run_serve(args)
    return {
        "metric": METRIC, "value": s["value"], "unit": "G disparity-evals/s", "n_gpus": 1,
        "steps": s["requests"], "warmup": 0, "ms_per_step": 1e3 / s["rigs_per_s"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u16", "data": "synthetic",
        "config": {"workload": s["workload"]}, **s,
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


def valu_peak():
    """Sustained VALU issue rate of v_add_u32, v_pk_add_u16 and v_bcnt_u32_b32
    (sister_amd/csrc/valu_peak.hip: 16 independent accumulators per lane, no
    memory traffic), against the spec-sheet 256 CUs x 64 lanes x 2.4 GHz."""
    import ctypes as C
    import torch  # noqa: F401  (binds the HIP runtime torch uses)
    path = os.path.join(ROOT, "sister_amd", "libvalu_peak.so")
    lib = C.CDLL(path)
    lib.valu_peak_run.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    out = {"spec_peak_lane_ops": VALU_PEAK}
    for op, name, per_inst in ((0, "v_add_u32", 1), (1, "v_pk_add_u16", 2), (2, "v_bcnt_u32_b32", 1)):
        ms, ops = C.c_double(), C.c_double()
        rc = lib.valu_peak_run(op, 256 * 16, 8192, C.byref(ms), C.byref(ops))
        if rc != 0:
            raise RuntimeError(f"valu_peak_run({op}) returned {rc}")
        rate = ops.value / (ms.value * 1e-3)
        inst_rate = rate / per_inst
        out[name] = {"ms": ms.value, "lane_ops_per_s": rate, "instr_lane_issues_per_s": inst_rate,
                     "frac_of_spec_issue": inst_rate / VALU_PEAK,
                     "implied_clock_GHz": inst_rate / (256 * 64) / 1e9}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--workload", default="rig1080", choices=["rig1080", "mixed32", "large", "serve"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="headline only (no mixed32 / serve blocks)")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-host", action="store_true", help="serve host buffers instead of device-resident ones")
    ap.add_argument("--hbm-gb", type=int, default=64)
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--valu-peak", action="store_true", help="VALU throughput probe (int32, packed u16, popcount)")
    ap.add_argument("--single", action="store_true", help="one eager run of the batch (counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only; --gpus must be 1 (rigs are independent, "
                                   "a multi-GPU run would shard rigs across processes)"}))
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.valu_peak:
        valu_peak()
        return
    if args.single:
        run_single(args)
        return
    if args.workload == "serve":
        print(json.dumps(serve_line(args)))
        return
    out = run(args)
    if args.workload == "rig1080" and not args.no_extra:
        out["mixed32"] = run_mixed(args)
        out["serve"] = run_serve(args, seconds=5.0)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
