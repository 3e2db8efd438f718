#!/usr/bin/env python3
"""Benchmark: list-mode TOF OSEM with 4-ray Siddon on one MI355X.

Default workload (--workload motion: the batched, motion-corrected list-mode
configuration): the 64-ring x 512-crystal scanner (radius 400 mm,
3.2 x 3.2 x 20 mm crystals), 250 M list-mode events (uint32 d1, d2, float32
TOF, u8 motion frame) in 8 rigid motion frames of uneven counts (uniform in
[5 M, 60 M], seeded, scaled to 250 M; rotation <= 3 deg, translation
<= 5 mm), each frame a stretch of the acquisition whose events are drawn
from the phantom (water cylinder r = 150 mm, length 180 mm, plus 6 spheres
of radius 4-14 mm at 5:1 contrast) as it lay during that frame; TOF 250 ps
FWHM truncated at 5 sigma; image 256 x 256 x 128 at 1.6 mm; 10 subsets.
The reconstruction moves every LOR back by its frame's transform inside the
projector; the sensitivity image is time-averaged over the motion.

One step = one full OSEM iteration: for each of the 10 subsets, the EM
update image of that subset's events (forward projection and ratio per
event, then backprojection) followed by the multiplicative image update.
Batched (the value): all motion frames' events of a subset in one launch
each of the EM-ratio kernel and the tile-planned backprojection (the Siddon
walks binned by image tiles, accumulated in LDS; plans built once in setup;
launch shapes and the persistent grid picked by the autotuner on subset 0,
--no-tune: defaults), the whole iteration replayed as one recorded graph
(--no-graph: launched call by call).  Events, plans, the detector LUT and
all images are resident in HBM before timing starts; the sensitivity image
(backprojection of all 536.9 M scanner LORs x 4 rays), the plans and the
tuning are made once in setup and timed separately.  Per-kernel times
(roofline) come from one more iteration run call by call with HIP events on
the context's stream.

Beside it, from the same starting image for the configuration's 3
iterations: the plain per-frame path (each motion frame's events of a
subset launched on their own, same kernels and per-frame tile plans, no
persistent grid, no graph) and the plain fused path (the one-kernel EM with
global atomics, per frame), both timed, and the relative L2 of their images
against the batched one; the EM update image and the forward projections
of an evenly strided sample of subset 0 against the oracle's FP32 and FP64
restatements; and the CPU baseline (the oracle's FP32 restatement of the
reference EM subset on 16 threads over a bounded sample).

--workload listmode: the 100 M-event configuration without motion
(200 x 200 x 128 at 2 mm), the same batched path, oracle sample and CPU
baseline.

--workload histogram: the CPU-runnable histogram configuration (24 rings x
256 crystals, R 300 mm, max ring difference 23; image 96 x 96 x 48 at 2 mm;
water cylinder r 100 mm x 90 mm with 4 hot spheres (r 5/8/11/15 mm, 4:1),
mu 0.0096/mm inside; measured histogram = Poisson(attenuated forward
projection scaled to 50 M counts); Siddon, no TOF, 5 iterations x 8
subsets), reconstructed end to end through the C++ host layer (pyyrtmi.OSEM:
sensitivity images, then reconstruct) and by the oracle's restatement of
the reference CPU path on the same data, timed beside it; prints the final
images' relative L2 difference.

--workload frames: 16 dynamic frames of uneven event counts (log-uniform in
[--frame-min, --frame-max], seeded) reconstructed (a) the plain way, frame
after frame with the fused EM kernel, (b) frame after frame through tile
plans, and (c) batched: all frames in one multi-frame image and one event
set (frame per event), tile plans, autotuned, graph-replayed.

--workload batch: 8 independent images of 128x128x64 @2 mm, 20M TOF
(350 ps) events each, one launch per subset for all eight.

--workload dd: the distance-driven dynamic configuration (see bench_dd).

Prints one JSON line (see the dict each workload returns).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [REPO, os.path.join(REPO, "yrt-pet_amd")]

HBM_PEAK_GBS = 8000.0
# Per-launch HBM traffic of the list-mode configurations, from two rocprofv3
# counter passes (FETCH_SIZE, WRITE_SIZE) scaled by a copy of known size;
# each file gives its commands.
TRAFFIC_PROFILES = {
    "motion": os.path.join(REPO, "profiles", "em_motion_tof_250M_pmc_traffic.md"),
    "listmode": os.path.join(REPO, "profiles", "em_listmode_tof_100M_pmc_traffic.md"),
    "dd": os.path.join(REPO, "profiles", "dd_psf_histo_16f_pmc.md"),
}


def profiled_traffic(path, kernel_label):
    """HBM bytes per launch of `kernel_label` (its part before the first
    space) from a committed counter summary, or None."""
    import re
    try:
        text = open(path).read()
    except OSError:
        return None
    short = kernel_label.split(" ")[0]
    for m in re.finditer(r"^- traffic `([^`]+)`: ([0-9.eE+-]+) B$", text, re.M):
        if m.group(1) == short:
            return float(m.group(2))
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=("motion", "listmode", "frames", "histogram", "dd",
                                           "batch", "autograd"),
                    default="motion")
    ap.add_argument("--dd-min", type=float, default=1e6, help="dd workload: min counts/frame")
    ap.add_argument("--dd-max", type=float, default=40e6, help="dd workload: max counts/frame")
    ap.add_argument("--dd-subsets", type=int, default=12)
    ap.add_argument("--ag-events", type=int, default=20_000_000,
                    help="autograd workload: events per image")
    ap.add_argument("--events", type=int, default=0,
                    help="list-mode events (0: 250 M with motion, 100 M without)")
    ap.add_argument("--subsets", type=int, default=10)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--lpe", type=int, default=4, help="lanes per event (untuned runs)")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--accum", choices=("tiled", "xcd", "direct"), default="tiled",
                    help="backprojection: tile plans with LDS accumulation, or global "
                         "atomics into per-XCD images / the target directly")
    ap.add_argument("--chunk", type=int, default=0, help="tile plan records per workgroup")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--tile-shape", type=int, default=0,
                    help="tile shape of untuned runs (yrtmi.capi.TILE_*)")
    ap.add_argument("--order", choices=("lor", "none"), default="lor",
                    help="project each subset's events in locality order (yrtmi_lor_order)")
    ap.add_argument("--no-graph", action="store_true", help="launch the iteration call by call")
    ap.add_argument("--frames", type=int, default=16)
    ap.add_argument("--frame-min", type=float, default=0.25e6)
    ap.add_argument("--frame-max", type=float, default=10e6)
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 256 MiB device copy first (PMC calibration)")
    ap.add_argument("--no-plain", action="store_true",
                    help="motion: skip the plain per-frame comparison runs")
    ap.add_argument("--no-plain-fused", action="store_true",
                    help="motion: skip the plain per-frame fused-kernel run")
    ap.add_argument("--no-oracle-sample", action="store_true",
                    help="skip the EM-update parity sample against the oracle")
    ap.add_argument("--oracle-events", type=int, default=1_000_000,
                    help="events of the oracle parity sample")
    ap.add_argument("--no-dropin", action="store_true",
                    help="motion workload: skip the OSEM_MI (reconstruct --gpu) leg")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def strided_sample(torch, n, k, dev):
    """k indices spread evenly over [0, n), first 0 and last n - 1, in exact
    integer arithmetic (a float32 linspace rounds indices near 1e8 up to n)."""
    k = max(1, min(int(k), int(n)))
    i = torch.arange(k, device=dev, dtype=torch.int64)
    return (i * (int(n) - 1)) // max(k - 1, 1)


def log(args, *a):
    if not args.quiet:
        print("[bench]", *a, file=sys.stderr, flush=True)


def scanner_and_image(motion=False):
    import yrtmi
    sc = yrtmi.Scanner("cfg2", 64 * 3.2, 3.2, 3.2, 20.0, 400.0, 512, 64)
    if motion:
        # the motion-corrected configuration: 256x256x128 @1.6 mm
        return sc, yrtmi.ImageParams(256, 256, 128, 1.6, 1.6, 1.6)
    im = yrtmi.ImageParams(200, 200, 128, 2.0, 2.0, 2.0)
    return sc, im


def phantom(args, im):
    from yrtmi import synth
    rng = np.random.default_rng(args.seed)
    spheres = synth.random_spheres_in_cylinder(rng, 150.0, 180.0,
                                               list(np.linspace(4.0, 14.0, 6)), 5.0)
    return synth.cylinder_spheres(im, 150.0, 180.0, 1.0, spheres, supersample=2)


def subset_major(torch, ev, n, S, dev):
    """Event i belongs to subset i mod S (BinIteratorChronologicalInterleaved);
    permute once so each subset is contiguous."""
    perm = torch.cat([torch.arange(s, n, S, device=dev) for s in range(S)])
    out = {k: ev[k][perm].contiguous() for k in ("d1", "d2", "tof", "motion") if k in ev}
    bounds = [0]
    for s in range(S):
        bounds.append(bounds[-1] + len(range(s, n, S)))
    return out, bounds


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus must be 1 (the events of a "
                                   "subset would shard across GPUs, with the tile images "
                                   "summed by an all-reduce; that path is not built)"}))
        return 2
    import torch
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calib_copy:
        # a plain copy of known size: 256 MiB read + 256 MiB written
        a = torch.ones(64 << 20, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    stream = torch.cuda.Stream()  # default priority; graphs need a non-null stream
    with torch.cuda.stream(stream):
        if args.workload in ("frames", "batch"):
            out = bench_frames(args, torch, dev, stream)
        elif args.workload == "histogram":
            out = bench_histogram(args, torch, dev, stream)
        elif args.workload == "dd":
            out = bench_dd(args, torch, dev, stream)
        elif args.workload == "autograd":
            out = bench_autograd(args, torch, dev, stream)
        else:
            out = bench_listmode(args, torch, dev, stream)
    print(json.dumps(out))
    return 0


def motion_frame_counts(args, nmf, total):
    """Events per motion frame: uniform in [5 M, 60 M] (seeded), scaled to
    the total; frames are contiguous stretches of the acquisition."""
    rng = np.random.default_rng(args.seed + 5)
    c = rng.uniform(5e6, 60e6, nmf)
    c = np.floor(c * (total / c.sum())).astype(np.int64)
    c[-1] += total - int(c.sum())
    return [int(v) for v in c]


def bench_listmode(args, torch, dev, stream):
    """List-mode TOF OSEM, 4-ray Siddon: the motion-corrected configuration
    (default) or the 100 M-event one (--workload listmode)."""
    import yrtmi
    from yrtmi import synth

    t_setup = time.time()
    motion = args.workload == "motion"
    sc, im = scanner_and_image(motion)
    act = phantom(args, im)
    log(args, f"phantom ready ({time.time() - t_setup:.1f}s)")
    n_events = args.events if args.events > 0 else (250_000_000 if motion else 100_000_000)
    S, n = args.subsets, n_events
    NMF = 8 if motion else 0
    xf, fcounts = None, None
    if motion:
        # 8 rigid motion frames (rotation <= 3 deg, translation <= 5 mm), each
        # a stretch of the acquisition whose events come from the object as
        # it lay then; the reconstruction moves every LOR back (p' = R p + t)
        xf = synth.rigid_transforms(np.random.default_rng(args.seed), NMF)
        fcounts = motion_frame_counts(args, NMF, n)
        parts = [synth.sample_events(sc, im, act, fcounts[f], seed=args.seed + 13 * f,
                                     tof_fwhm_ps=250.0, device=dev, motion_xf=xf[f])
                 for f in range(NMF)]
        ev = {k: torch.cat([p_[k] for p_ in parts]) for k in ("d1", "d2", "tof")}
        ev["motion"] = torch.cat([torch.full((c,), f, dtype=torch.uint8, device=dev)
                                  for f, c in enumerate(fcounts)])
        del parts
    else:
        ev = synth.sample_events(sc, im, act, n, seed=args.seed, tof_fwhm_ps=250.0, device=dev)
    ev, bounds = subset_major(torch, ev, n, S, dev)
    d1, d2, tof = ev["d1"], ev["d2"], ev["tof"]
    mot = ev.get("motion")
    del ev
    torch.cuda.synchronize()
    log(args, f"{n} events drawn ({time.time() - t_setup:.1f}s)"
              + (f"; motion frames {fcounts}" if motion else ""))

    ctx = yrtmi.Context(sc, stream=stream)
    if motion:
        ctx.upload_motion(xf)
    pp = yrtmi.ProjectorParams(sc, numRays=4)
    pp.addTOF(250.0, 5)
    proj = ctx.projector(pp)
    tiled = args.accum == "tiled"
    proj.set_launch(args.lpe, args.block, args.grid)
    proj.set_accumulation(args.accum == "xcd")
    # sensitivity image: all scanner LORs (uniform histogram), no TOF
    psens = ctx.projector(pp, with_tof=False)
    psens.set_launch(args.lpe, args.block, args.grid)
    psens.set_accumulation(args.accum == "xcd")
    sens = torch.zeros(im.shape, device=dev)
    torch.cuda.synchronize()
    t0 = time.time()
    SS = 16
    for s in range(SS):
        Ls = yrtmi.LorSet.histo_subset(ctx, SS, s)
        if tiled:
            pl = ctx.plan(psens, im, Ls, args.chunk, tile_shape=yrtmi.capi.TILE_32x32x8,
                          single_use=True)
            ctx.sensitivity(psens, im, sens, Ls, plan=pl)
            del pl
        else:
            ctx.sensitivity(psens, im, sens, Ls)
    if motion:
        # the sensitivity image time-averaged over the motion frames, each
        # weighted by its share of the acquisition, as the reconstruct
        # command does with the motion file's frame durations
        moved = torch.zeros_like(sens)
        ctx.time_average_move(im, sens, moved, xf,
                              np.asarray(fcounts, np.float64) / n)
        sens = moved
    ctx.synchronize()
    sens_s = time.time() - t0
    n_lors = sum(yrtmi.LorSet.histo_subset(ctx, SS, s).n for s in range(SS))
    log(args, f"sensitivity: {n_lors} LORs x 4 rays in {sens_s:.2f}s")
    # hard threshold (voxels <= 1 -> 0), list-mode scaling 1/S, mask, init
    ctx.threshold(sens, sens.clone(), im.nvox, 1.0, 0.0, 0.0, 1.0, 0.0)
    sens_thr = sens.clone() if motion and not args.no_dropin else None
    sens *= float(np.float32(1.0) / np.float32(S))
    x = torch.full(im.shape, 0.125, device=dev)
    ctx.threshold(x, sens, im.nvox, 0.0, 0.0, 0.0, 1.0, 0.0)
    x0 = x.clone()
    u = torch.zeros(im.shape, device=dev)

    def lors_of(a, b):
        return yrtmi.LorSet.pairs(d1[a:b], d2[a:b], tof=tof[a:b],
                                  motion_frame=None if mot is None else mot[a:b])

    sub = [lors_of(bounds[s], bounds[s + 1]) for s in range(S)]
    order_s = 0.0
    if args.order == "lor":
        # each subset's events in locality order (the EM update does not
        # depend on it); a setup step like the plans
        ctx.synchronize()
        t0 = time.time()
        sub = [L.permuted(ctx, ctx.lor_order(im, L)) for L in sub]
        ctx.synchronize()
        order_s = time.time() - t0
        log(args, f"locality order of the subsets: {order_s:.2f}s")
    # algorithmic work: voxel segments per subset (instrumented count)
    segs = [ctx.count_segments(proj, im, L) for L in sub]
    log(args, f"segments/event {sum(segs) / n:.1f}; setup {time.time() - t_setup:.1f}s")
    plans, plan_s, plan_bytes, plan_records, tune, tune_s = [], 0.0, 0, 0, None, 0.0
    shape = args.tile_shape
    vbuf = None
    if tiled:
        pers = 0
        if not args.no_tune:
            t0 = time.time()
            tune = ctx.autotune(proj, im, sub[0], x)
            tune_s = time.time() - t0
            shape, pers = tune["tile_shape"], tune["persistent"]
            log(args, f"autotune ({tune_s:.2f}s): {tune}")
        ctx.synchronize()
        t0 = time.time()
        plans = [ctx.plan(proj, im, L, args.chunk, tile_shape=shape, persistent=pers)
                 for L in sub]
        ctx.synchronize()
        plan_s = time.time() - t0
        infos = [p_.info() for p_ in plans]
        plan_bytes = sum(i["bytes"] for i in infos)
        plan_records = sum(i["records"] for i in infos)
        vbuf = torch.empty(max(L.n for L in sub), device=dev)
        log(args, f"tile plans: {plan_records} records ({plan_records / (4 * n):.2f} per ray), "
                  f"{plan_bytes / 2**30:.1f} GiB, {plan_s:.2f}s")

    def iteration(evs=None):
        for s in range(S):
            ctx.fill(u, 0.0)
            if evs is not None:
                evs[s][0].record()
            if tiled:
                ctx.em_ratio(proj, im, x, sub[s], vbuf)
                if evs is not None:
                    evs[s][1].record()
                ctx.back(proj, im, u, sub[s], vbuf, plan=plans[s])
            else:
                ctx.em_subset(proj, im, x, u, sub[s])
            if evs is not None:
                evs[s][2].record()
            ctx.em_update(x, u, sens, im.nvox)

    for _ in range(args.warmup):
        iteration()
    graph = None
    if not args.no_graph:
        ctx.synchronize()
        with ctx.graph() as graph:
            iteration()
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        if graph is not None:
            graph.launch()
        else:
            iteration()
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # per-kernel times: one more iteration call by call, HIP events on the
    # context's stream (torch's current stream here)
    evs = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(S)]
    iteration(evs)
    ctx.synchronize()
    s_per_iter = wall / args.steps
    ev_proj_per_s = 2.0 * n * args.steps / wall
    # Algorithmic bytes per launch (R_ev = 12 B/event: d1, d2, tof, +1 B for
    # the motion frame; an atomic counts as an 8 B read-modify-write):
    #   forward + ratio:  R_ev + 4 B per voxel segment + 4 B ratio stored
    #   backprojection:   R_ev + 4 B value + 8 B per voxel segment
    #   fused EM:         R_ev + 12 B per voxel segment
    r_ev = 13.0 if motion else 12.0
    nev = [bounds[s + 1] - bounds[s] for s in range(S)]
    kernels = []

    def roof(name, idx, bytes_per_subset):
        ms = [evs[s][idx[0]].elapsed_time(evs[s][idx[1]]) for s in range(S)]
        avg_ms, avg_b = float(np.mean(ms)), float(np.mean(bytes_per_subset))
        ach = avg_b / (avg_ms * 1e-3) / 1e9
        kernels.append({"kernel": name, "avg_launch_ms": avg_ms, "alg_bytes_per_launch": avg_b,
                        "achieved": ach, "frac": ach / HBM_PEAK_GBS})

    lpe = tune["lanes_per_event"] if tune else args.lpe
    if tiled:
        roof(f"kEmRatio<TOF,LPE={lpe}> (forward + ratio)", (0, 1),
             [r_ev * nev[s] + 4.0 * segs[s] + 4.0 * nev[s] for s in range(S)])
        roof(f"kTileBack<TOF,shape={shape}> (tile-binned backprojection)", (1, 2),
             [r_ev * nev[s] + 4.0 * nev[s] + 8.0 * segs[s] for s in range(S)])
    else:
        roof(f"kEvents<MODE_EM,Siddon,TOF,LPE={lpe}> (fused EM)", (0, 2),
             [r_ev * nev[s] + 12.0 * segs[s] for s in range(S)])
    dom = max(kernels, key=lambda k: k["avg_launch_ms"])
    log(args, f"{s_per_iter:.3f} s/iter ({'graph' if graph else 'calls'}); " + "; ".join(
        f"{k['kernel']}: {k['avg_launch_ms']:.2f} ms, {k['achieved']:.0f} GB/s alg" for k in kernels))
    # measured traffic applies to the profiled configuration only
    traffic, traffic_src = None, None
    prof = TRAFFIC_PROFILES.get("motion" if motion else "listmode")
    if (tiled and n == (250_000_000 if motion else 100_000_000) and S == 10
            and args.order == "lor" and not args.no_tune):
        traffic = profiled_traffic(prof, dom["kernel"])
        if traffic is not None:
            traffic_src = os.path.relpath(prof, REPO)

    # ---- parity: the configuration's 3 iterations from the same start,
    # batched here, the plain per-frame paths below; and the EM update of a
    # sample of subset 0 against the oracle
    iters_cmp = 3
    xb = None
    parity = {}
    if tiled and (not args.no_plain or not args.no_oracle_sample or sens_thr is not None):
        x.copy_(x0)
        for _ in range(iters_cmp):
            if graph is not None:
                graph.launch()
            else:
                iteration()
        ctx.synchronize()
        xb = x.clone()
    if not args.no_oracle_sample:
        parity["oracle_sample"] = oracle_sample(args, torch, ctx, proj, sc, im, pp, d1, d2, tof,
                                                mot, xf, bounds, xb if xb is not None else x,
                                                shape, dev)
    plain = None
    if motion and tiled and not args.no_plain:
        del graph
        graph = None
        plans = []
        plain = plain_per_frame(args, torch, ctx, proj, im, d1, d2, tof, mot, bounds, fcounts,
                                x, x0, u, sens, shape, vbuf, xb, iters_cmp)
        log(args, "plain per-frame: " + json.dumps(plain))
    dropin = None
    if motion and tiled and sens_thr is not None:
        del graph
        graph = None
        plans = []
        dropin = dropin_osem(args, torch, d1, d2, tof, mot, xf, sens_thr, S, xb, iters_cmp)
        log(args, "drop-in OSEM_MI: " + json.dumps(dropin))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, sc, im, pp, d1, d2, tof, bounds, x0, mot, xf)
    wl = ((f"list-mode TOF OSEM with LOR motion ({NMF} rigid frames of uneven counts), "
           f"4-ray Siddon, 64x512 scanner R400, {n / 1e6:g}M events, "
           f"256x256x128 @1.6mm, {S} subsets") if motion else
          (f"list-mode TOF OSEM, 4-ray Siddon, 64x512 scanner R400, "
           f"{n / 1e6:g}M events, 200x200x128 @2mm, {S} subsets"))
    out = {
        "metric": "fwd+back event-projections/s, s/OSEM iter; batched vs plain at <=1e-4 rel L2",
        "value": ev_proj_per_s,
        "unit": "event-projections/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_per_iter * 1e3,
        "s_per_osem_iter": s_per_iter,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": ("synthetic (seeded phantom; TOF list-mode events sampled physically, "
                 "each motion frame's from the moved object)" if motion else
                 "synthetic (seeded phantom, physically sampled TOF list-mode events)"),
        "config": {"workload": wl,
                   "motion_frames": NMF, "motion_frame_events": fcounts,
                   "events": n, "subsets": S, "rays": 4, "tof_fwhm_ps": 250.0,
                   "tof_n_std": 5, "image": [im.nx, im.ny, im.nz], "voxel_mm": float(im.vx),
                   "lanes_per_event": lpe, "accumulation": args.accum,
                   "event_order": args.order, "batched": "all motion frames in one launch "
                   "per subset" if motion else "one launch per subset",
                   "graph": not args.no_graph, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": dom["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["frac"], "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": dom["kernel"], "avg_launch_ms": dom["avg_launch_ms"],
                     "alg_bytes_per_launch": dom["alg_bytes_per_launch"],
                     "segments_per_event": sum(segs) / n},
        "kernels": kernels,
        "tuning": tune,
        "tuning_s": tune_s,
        "sensitivity_s": sens_s,
        "sensitivity_lors": n_lors,
        "order_s": order_s,
        "plan_s": plan_s,
        "plan_bytes": plan_bytes,
        "plan_records": plan_records,
        "parity": parity,
        "cpu_baseline": cpu,
    }
    if plain is not None:
        out["plain"] = plain
        out["plain_value"] = plain["per_frame_tiled_value"]
        out["speedup_vs_plain"] = plain["per_frame_tiled_s"] / s_per_iter
        if "per_frame_fused_s" in plain:
            out["speedup_vs_plain_fused"] = plain["per_frame_fused_s"] / s_per_iter
        out["parity"]["batched_vs_plain_rel_l2"] = plain["rel_l2_vs_batched"]
    if dropin is not None:
        out["dropin"] = dropin
    return out


def dropin_osem(args, torch, d1, d2, tof, mot, xf, sens_thr, S, xb, iters):
    """The same reconstruction through the C++ OSEM (OSEM_MI: the reconstruct
    command's --gpu path): events aliased on the device in acquisition order
    (event i in subset i mod S), the bench's thresholded sensitivity image,
    graph-replayed subsets, autotuned launch shape, plans within 90 % of the
    free device memory.  Its time per iteration is that of the graph-replayed
    iterations, which OSEM_MI times itself (timeIterations: the device
    synchronized after each); `setup_s` is the rest of the call (plans,
    locality order, autotune).  Its image after `iters` iterations is
    compared with the batched one (EM from a uniform start does not depend on
    the start's value)."""
    import pyyrtmi as Y
    n = d1.numel()
    # back to acquisition order (the bench keeps each subset contiguous)
    perm = torch.cat([torch.arange(s, n, S, device=d1.device) for s in range(S)])
    acq = {}
    for k, v in (("d1", d1), ("d2", d2), ("tof", tof), ("motion", mot)):
        a = torch.empty_like(v)
        a[perm] = v
        acq[k] = a
    del perm
    ysc = Y.Scanner("cfg2", 64 * 3.2, 3.2, 3.2, 20.0, 400.0, 512, 64, 1, 63, 1, 1)
    yim = Y.ImageParams.fromParams(256, 256, 128, 1.6, 1.6, 1.6)
    osem = Y.createOSEM(ysc)
    octx = osem.getContext()
    lm = Y.ListModeDeviceAlias(octx, n, acq["d1"], acq["d2"], tof=acq["tof"],
                               motion=acq["motion"])
    lm.setMotionTransforms(np.ascontiguousarray(xf, np.float32).reshape(-1))
    osem.setImageParams(yim)
    osem.setNumRays(4)
    osem.addTOF(250.0, 5)
    osem.num_OSEM_subsets = S
    osem.setDataInput(lm)
    simg = Y.ImageDeviceAlias(octx, yim, sens_thr)
    osem.setSensitivityImages([simg])
    osem.num_MLEM_iterations = iters
    osem.timeIterations = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    img = osem.reconstruct()
    torch.cuda.synchronize()
    call_s = time.perf_counter() - t0
    its = list(osem.iterationSeconds())
    x = img.transferToHostMemory().reshape(tuple(xb.shape))
    # the graph-replayed iterations (after the first, call by call, and the
    # second, recorded then launched)
    replay = its[2:] if len(its) > 2 else its[-1:]
    per_iter = float(np.mean(replay))
    out = {"s_per_osem_iter": per_iter, "value": 2.0 * n / per_iter,
           "iteration_s": its, "call_s": call_s, "setup_s": call_s - sum(its),
           "planned_subsets": osem.numPlannedSubsets(),
           "rel_l2_vs_batched": float(np.linalg.norm(x - xb.cpu().numpy()) /
                                      np.linalg.norm(xb.cpu().numpy()))}
    del lm, osem, acq
    return out


def rel_l2_t(a, b):
    a = a.double().reshape(-1)
    b = b.double().reshape(-1)
    nb = float(b.norm())
    return float((a - b).norm()) / (nb if nb > 0 else 1.0)


def plain_per_frame(args, torch, ctx, proj, im, d1, d2, tof, mot, bounds, fcounts, x, x0, u,
                    sens, shape, vbuf, xb, iters):
    """The plain per-frame path: every subset's events split by motion frame,
    each frame projected by its own launches (EM ratio + tile-planned
    backprojection with the same kernels, plans per frame and subset, no
    persistent grid, no graph), then the fused one-kernel EM per frame; each
    run for `iters` iterations from the same start as the batched run, timed,
    and compared with it."""
    import yrtmi
    S, F = len(bounds) - 1, len(fcounts)
    cum = np.concatenate([[0], np.cumsum(fcounts)])
    per = []   # per[s][f]: the LORs of subset s from motion frame f
    ctx.synchronize()
    t0 = time.time()
    for s in range(S):
        row, off = [], bounds[s]
        for f in range(F):
            # acquisition events i = s (mod S) in [cum[f], cum[f+1]) sit
            # contiguously in subset s
            first = int(cum[f]) + ((s - int(cum[f])) % S)
            cnt = max(0, (int(cum[f + 1]) - first + S - 1) // S)
            L = yrtmi.LorSet.pairs(d1[off:off + cnt], d2[off:off + cnt], tof=tof[off:off + cnt],
                                   motion_frame=mot[off:off + cnt])
            if args.order == "lor" and cnt > 1:
                L = L.permuted(ctx, ctx.lor_order(im, L))
            row.append(L)
            off += cnt
        assert off == bounds[s + 1], "per-frame split does not cover the subset"
        per.append(row)
    fplans = [[ctx.plan(proj, im, per[s][f], args.chunk, tile_shape=shape, persistent=0)
               if per[s][f].n > 0 else None for f in range(F)] for s in range(S)]
    ctx.synchronize()
    res = {"per_frame_plan_s": time.time() - t0,
           "per_frame_plan_bytes": sum(p_.info()["bytes"] for r in fplans for p_ in r if p_)}

    def it_tiled():
        for s in range(S):
            ctx.fill(u, 0.0)
            for f in range(F):
                if per[s][f].n > 0:
                    ctx.em_ratio(proj, im, x, per[s][f], vbuf)
                    ctx.back(proj, im, u, per[s][f], vbuf, plan=fplans[s][f])
            ctx.em_update(x, u, sens, im.nvox)

    def it_fused():
        for s in range(S):
            ctx.fill(u, 0.0)
            for f in range(F):
                if per[s][f].n > 0:
                    ctx.em_subset(proj, im, x, u, per[s][f])
            ctx.em_update(x, u, sens, im.nvox)

    def run(fn, name):
        x.copy_(x0)
        fn()                       # warmup iteration
        x.copy_(x0)
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(iters):
            fn()
        ctx.synchronize()
        res[f"per_frame_{name}_s"] = (time.perf_counter() - t) / iters
        res[f"per_frame_{name}_value"] = 2.0 * int(bounds[-1]) / res[f"per_frame_{name}_s"]
        if xb is not None:
            res[f"rel_l2_{name}_vs_batched"] = rel_l2_t(x, xb)

    run(it_tiled, "tiled")
    res["rel_l2_vs_batched"] = res.get("rel_l2_tiled_vs_batched")
    del fplans
    if not args.no_plain_fused:
        proj.set_accumulation(True)      # the fused kernel's per-XCD images
        run(it_fused, "fused")
        proj.set_accumulation(False)
    res["iterations"] = iters
    res["note"] = ("per_frame_tiled: the same EM-ratio and tile-backprojection kernels launched "
                   "per motion frame and subset, plans per frame, no persistent grid, no graph; "
                   "per_frame_fused: the one-kernel EM with global atomics per frame")
    return res


def oracle_sample(args, torch, ctx, proj, sc, im, pp, d1, d2, tof, mot, xf, bounds, x, shape,
                  dev):
    """The EM update image and the forward projections of an evenly strided
    sample of subset 0 (every motion frame represented), GPU (EM ratio +
    tile-planned backprojection, the batched kernels) against the oracle's
    FP32 and FP64 restatements of the reference CPU path."""
    import oracle as O
    import yrtmi
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_image, oracle_proj, oracle_scanner, rel_l2
    k = min(args.oracle_events, bounds[1])
    idx = strided_sample(torch, bounds[1], k, dev)
    a, b, t = d1[idx].contiguous(), d2[idx].contiguous(), tof[idx].contiguous()
    m = None if mot is None else mot[idx].contiguous()
    L = yrtmi.LorSet.pairs(a, b, tof=t, motion_frame=m)
    y = torch.empty(k, device=dev)
    ctx.forward(proj, im, x, L, y)
    v = torch.empty(k, device=dev)
    ug = torch.zeros(im.shape, device=dev)
    pl = ctx.plan(proj, im, L, tile_shape=shape, single_use=True)
    ctx.em_ratio(proj, im, x, L, v)
    ctx.back(proj, im, ug, L, v, plan=pl)
    ctx.synchronize()
    nthreads = max(1, min(16, os.cpu_count() or 1))
    osc = oracle_scanner(sc)
    pos, ori = O.lut(osc)
    OL = O.lors_from_pairs(pos, ori, a.cpu().numpy().astype(np.uint32),
                           b.cpu().numpy().astype(np.uint32), tof=t.cpu().numpy(),
                           motion_frame=None if m is None else m.cpu().numpy(),
                           motion_xf=xf)
    po, oi = oracle_proj(pp), oracle_image(im)
    xh = x.cpu().numpy()
    t0 = time.perf_counter()
    y32 = O.forward(oi, xh, po, OL, nthreads=nthreads)
    y64 = O.forward(oi, xh.astype(np.float64), po, OL, dtype=np.float64, nthreads=nthreads)
    u32 = O.em_subset(oi, xh, np.zeros(im.shape, np.float32), po, OL, nthreads=nthreads)
    u64 = O.em_subset(oi, xh.astype(np.float64), np.zeros(im.shape, np.float64), po, OL,
                      dtype=np.float64, nthreads=nthreads)
    dt = time.perf_counter() - t0
    yg, ugh = y.cpu().numpy(), ug.cpu().numpy()
    res = {"events": k,
           "sample": f"{k} events evenly strided over subset 0, image after 3 batched "
                     f"iterations",
           "fwd_rel_l2_gpu_vs_oracle_f32": rel_l2(yg, y32),
           "fwd_rel_l2_gpu_vs_oracle_f64": rel_l2(yg, y64),
           "fwd_rel_l2_oracle_f32_vs_f64": rel_l2(y32, y64),
           "em_rel_l2_gpu_vs_oracle_f32": rel_l2(ugh, u32),
           "em_rel_l2_gpu_vs_oracle_f64": rel_l2(ugh, u64),
           "em_rel_l2_oracle_f32_vs_f64": rel_l2(u32, u64),
           "oracle_s": dt}
    log(args, "oracle sample: " + json.dumps(res))
    return res


def bench_frames(args, torch, dev, stream):
    """Batched dynamic frames vs the plain per-frame path (see docstring)."""
    import yrtmi
    from yrtmi import synth

    sc, im1 = scanner_and_image()
    # batch: 8 independent images of 128x128x64 @2 mm, 20M TOF (350 ps) events each
    batch = args.workload == "batch"
    tof_ps = 350.0 if batch else 250.0
    if batch:
        im1 = yrtmi.ImageParams(128, 128, 64, 2.0, 2.0, 2.0)
    F, S = (8 if batch else args.frames), args.subsets
    rng = np.random.default_rng(args.seed + 1)
    counts = np.exp(rng.uniform(math.log(args.frame_min), math.log(args.frame_max), F))
    counts = [20_000_000] * F if batch else [int(c) for c in counts]
    act = phantom(args, im1)
    t0 = time.time()
    evs = [synth.sample_events(sc, im1, act, c, seed=args.seed + 7 * f, tof_fwhm_ps=tof_ps,
                               device=dev) for f, c in enumerate(counts)]
    n = sum(counts)
    log(args, f"{F} frames, {n} events ({min(counts)}..{max(counts)}) in {time.time() - t0:.1f}s")
    ctx = yrtmi.Context(sc, stream=stream)
    pp = yrtmi.ProjectorParams(sc, numRays=4)
    pp.addTOF(tof_ps, 5)
    proj = ctx.projector(pp)
    proj.set_launch(args.lpe)
    sens = torch.full(im1.shape, 1.0, device=dev)  # a sensitivity of ones: EM-loop timing

    # per-frame subsets, contiguous
    per = []
    for f in range(F):
        e, b = subset_major(torch, evs[f], counts[f], S, dev)
        per.append([yrtmi.LorSet.pairs(e["d1"][b[s]:b[s + 1]], e["d2"][b[s]:b[s + 1]],
                                       tof=e["tof"][b[s]:b[s + 1]]) for s in range(S)])
    imF = yrtmi.ImageParams(im1.nx, im1.ny, im1.nz, im1.vx, im1.vy, im1.vz, nt=F)
    xF = torch.full(imF.shape, 0.125, device=dev)
    uF = torch.zeros(imF.shape, device=dev)
    x1 = [xF[f:f + 1] for f in range(F)]
    u1 = [uF[f:f + 1] for f in range(F)]

    def plain(planned=None):
        for f in range(F):
            for s in range(S):
                ctx.fill(u1[f], 0.0)
                ctx.em_subset(proj, im1, x1[f], u1[f], per[f][s],
                              plan=None if planned is None else planned[f][s])
                ctx.em_update(x1[f], u1[f], sens, im1.nvox)

    def timed(fn, steps):
        fn()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t) / steps

    res = {}
    proj.set_accumulation(True)
    res["plain_fused_s"] = timed(plain, args.steps)
    log(args, f"plain (per frame, fused EM): {res['plain_fused_s']:.3f} s/iter")
    planned = [[ctx.plan(proj, im1, per[f][s]) for s in range(S)] for f in range(F)]
    res["plain_tiled_s"] = timed(lambda: plain(planned), args.steps)
    log(args, f"per frame, tile plans: {res['plain_tiled_s']:.3f} s/iter")
    del planned
    # batched: one event set per subset over all frames, frame per event
    cat = lambda key, s: torch.cat([getattr(per[f][s], key) for f in range(F)]).contiguous()
    bsub = []
    for s in range(S):
        fr = torch.cat([torch.full((per[f][s].n,), f, dtype=torch.int32, device=dev)
                        for f in range(F)])
        Lb = yrtmi.LorSet.pairs(cat("d1", s), cat("d2", s), tof=cat("tof", s), frame=fr)
        if args.order == "lor":
            Lb = Lb.permuted(ctx, ctx.lor_order(imF, Lb))  # frame-major locality order
        bsub.append(Lb)
    tune = ctx.autotune(proj, imF, bsub[0], xF)
    log(args, f"autotune: {tune}")
    bplans = [ctx.plan(proj, imF, L, tile_shape=tune["tile_shape"],
                       persistent=tune["persistent"]) for L in bsub]

    def batched():
        for s in range(S):
            ctx.fill(uF, 0.0)
            ctx.em_subset(proj, imF, xF, uF, bsub[s], plan=bplans[s])
            ctx.em_update(xF, uF, sens, im1.nvox)

    batched()
    ctx.synchronize()
    with ctx.graph() as g:
        batched()
    res["batched_s"] = timed(g.launch, args.steps)
    log(args, f"batched (one launch per subset for all frames, graph): {res['batched_s']:.3f} s/iter")
    value = 2.0 * n / res["batched_s"]
    return {
        "metric": "fwd+back event-projections/s, batched dynamic frames vs plain per-frame",
        "value": value,
        "unit": "event-projections/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": res["batched_s"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded phantom, physically sampled TOF list-mode events)",
        "config": {"workload": (f"batch of {F} independent images, 20M events each, "
                                f"4-ray Siddon TOF 350 ps, 128x128x64 @2mm, {S} subsets")
                   if batch else (f"{F} dynamic frames, log-uniform [{args.frame_min:g}, "
                                  f"{args.frame_max:g}] events, 4-ray Siddon TOF 250 ps, "
                                  f"200x200x128 @2mm, {S} subsets"),
                   "frame_events": counts, "events": n},
        "plain_fused_value": 2.0 * n / res["plain_fused_s"],
        "plain_tiled_value": 2.0 * n / res["plain_tiled_s"],
        "speedup_vs_plain_fused": res["plain_fused_s"] / res["batched_s"],
        "speedup_vs_per_frame_tiled": res["plain_tiled_s"] / res["batched_s"],
        "seconds": res, "tuning": tune,
    }


def bench_dd(args, torch, dev, stream):
    """Batched distance-driven histogram OSEM with projection-space PSF over
    16 dynamic frames (the 48-ring configuration).  Each frame is a
    Histogram3D of its sampled counts; per subset the nonzero bins of every
    frame (the reference GPU EM skips zero bins) are projected:
      plain       frame after frame, the fused one-kernel EM (global atomics);
      per frame   frame after frame, the same tile-planned kernels as batched
                  (plans per frame and subset, no persistent grid, no graph);
      batched     all frames' bins of a subset in one launch of each kernel
                  (frame per bin, frame-major locality order), autotuned tile
                  shape and persistent grid, the iteration replayed as one
                  graph (the value).
    Each from the same start for the configuration's 4 iterations; the
    images of the plain paths are compared with the batched one."""
    import yrtmi
    from yrtmi import synth

    t_setup = time.time()
    # 48 rings x 384 crystals, R 350 mm, 4 x 4 x 20 mm crystals
    sc = yrtmi.Scanner("cfg3", 192.0, 4.0, 4.0, 20.0, 350.0, 384, 48)
    im1 = yrtmi.ImageParams(128, 128, 96, 2.5, 2.5, 2.5)
    F, S, iters = 16, args.dd_subsets, 4
    rng = np.random.default_rng(args.seed + 3)
    counts = [int(c) for c in np.exp(rng.uniform(math.log(args.dd_min), math.log(args.dd_max), F))]
    # activity: water cylinder (1.0) and 5 spheres whose uptake follows a
    # one-tissue compartment curve 4 (1 - exp(-k2 t)) / (1 - exp(-k2 T)),
    # k2 ~ U[0.01, 0.3] / s, frames at t = 60 s (f + 1)
    spheres = synth.random_spheres_in_cylinder(rng, 120.0, 160.0, [6.0, 9.0, 12.0, 15.0, 18.0], 4.0)
    k2 = rng.uniform(0.01, 0.3, len(spheres))
    T = 60.0 * F
    ctx = yrtmi.Context(sc, stream=stream)
    nbins = yrtmi.LorSet.histo_subset(ctx, 1, 0).n
    hist = []
    for f in range(F):
        t = 60.0 * (f + 1)
        sph = [(cx, cy, cz, r, 1.0 + 3.0 * (1 - math.exp(-k * t)) / (1 - math.exp(-k * T)))
               for (cx, cy, cz, r, _), k in zip(spheres, k2)]
        act = synth.cylinder_spheres(im1, 120.0, 160.0, 1.0, sph, supersample=1)
        ev = synth.sample_events(sc, im1, act, counts[f], seed=args.seed + 11 * f, device=dev)
        b = ctx.histo_pair_bins(ev["d1"], ev["d2"]).long()
        hist.append(torch.bincount(b, minlength=nbins).float())
        del ev, b
    n_counts = sum(counts)
    pp = yrtmi.ProjectorParams(sc, projectorType="DD")
    k, s_step, ks = synth.gaussian_psf_row(2.0, k_spacing=0.25, num_s=1, s_step=1000.0)
    pp.psf_kernels, pp.psf_s_step, pp.psf_k_spacing = k, s_step, ks
    proj = ctx.projector(pp)
    sens = torch.full(im1.shape, 1.0, device=dev)  # a sensitivity of ones: EM-loop timing
    # per frame and subset: the nonzero bins and their counts
    per = [[ctx.histo_subset_nonzero(hist[f], S, s) for s in range(S)] for f in range(F)]
    del hist
    n = sum(per[f][s][0].numel() for f in range(F) for s in range(S))
    imF = yrtmi.ImageParams(im1.nx, im1.ny, im1.nz, im1.vx, im1.vy, im1.vz, nt=F)
    x0 = torch.full(imF.shape, 0.125, device=dev)
    xF = x0.clone()
    uF = torch.zeros(imF.shape, device=dev)
    x1 = [xF[f:f + 1] for f in range(F)]
    u1 = [uF[f:f + 1] for f in range(F)]
    Lf = [[yrtmi.LorSet.bins_(per[f][s][0]) for s in range(S)] for f in range(F)]
    # footprint voxels (the DD work unit), exactly
    fp = [sum(ctx.count_segments(proj, im1, Lf[f][s]) for f in range(F)) for s in range(S)]
    log(args, f"DD: {F} frames, {n_counts} counts in {n} nonzero bins "
              f"({min(counts)}..{max(counts)} counts/frame), {sum(fp) / n:.0f} footprint voxels "
              f"per bin; setup {time.time() - t_setup:.1f}s")

    def timed(fn, steps, reset=True):
        if reset:
            xF.copy_(x0)
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t) / steps

    res, rel = {}, {}
    # batched: one LOR set per subset over all frames, frame per bin
    bsub, bmeas = [], []
    for s in range(S):
        fr = torch.cat([torch.full((per[f][s][0].numel(),), f, dtype=torch.int32, device=dev)
                        for f in range(F)])
        Lb = yrtmi.LorSet.bins_(torch.cat([per[f][s][0] for f in range(F)]).contiguous(),
                                frame=fr)
        mb = torch.cat([per[f][s][1] for f in range(F)]).contiguous()
        if args.order == "lor":
            perm = ctx.lor_order(imF, Lb)  # frame-major locality order
            Lb, mb = Lb.permuted(ctx, perm), ctx.permute(mb, perm)
        bsub.append(Lb)
        bmeas.append(mb)
    t0 = time.time()
    tune = ctx.autotune(proj, imF, bsub[0], xF)
    tune_s = time.time() - t0
    log(args, f"autotune ({tune_s:.1f}s): {tune}")
    ctx.synchronize()
    t0 = time.time()
    plans = [ctx.plan(proj, imF, bsub[s], tile_shape=tune["tile_shape"],
                      persistent=tune["persistent"]) for s in range(S)]
    ctx.synchronize()
    res["plan_s"] = time.time() - t0
    plan_bytes = sum(p_.info()["bytes"] for p_ in plans)
    plan_records = sum(p_.info()["records"] for p_ in plans)
    vbuf = torch.empty(max(L.n for L in bsub), device=dev)

    def batched(evs=None):
        for s in range(S):
            ctx.fill(uF, 0.0)
            if evs is not None:
                evs[s][0].record()
            ctx.em_ratio(proj, imF, xF, bsub[s], vbuf, meas=bmeas[s], plan=plans[s])
            if evs is not None:
                evs[s][1].record()
            ctx.back(proj, imF, uF, bsub[s], vbuf, plan=plans[s])
            if evs is not None:
                evs[s][2].record()
            ctx.em_update(xF, uF, sens, im1.nvox)

    for _ in range(args.warmup):
        batched()
    ctx.synchronize()
    if args.no_graph:
        g = None
        step = batched
    else:
        with ctx.graph() as g:
            batched()
        step = g.launch
    res["batched_s"] = timed(step, args.steps, reset=False)
    # the configuration's iterations from the start, for the comparisons
    xF.copy_(x0)
    for _ in range(iters):
        step()
    ctx.synchronize()
    xb = xF.clone()
    evs = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(S)]
    batched(evs)
    ctx.synchronize()
    log(args, f"DD batched (one launch per subset, tile plans {plan_records / n:.1f} records/bin, "
              f"{plan_bytes / 2**30:.2f} GiB built in {res['plan_s']:.2f}s, graph): "
              f"{res['batched_s']:.3f} s/iter")
    # algorithmic bytes: R = 12 B per bin (bin id, frame, count); the ratio
    # kernel reads 4 B per footprint voxel and stores a ratio, the tile
    # backprojection reads the ratio and adds 8 B per footprint voxel
    kernels = []
    nb_s = [bsub[s].n for s in range(S)]
    for name, idx, byt in (("kEmRatioDD (forward + ratio)", (0, 1),
                            [12.0 * nb_s[s] + 4.0 * fp[s] + 4.0 * nb_s[s] for s in range(S)]),
                           (f"kTileBack<DD,shape={tune['tile_shape']}> (tile-binned "
                            "backprojection)", (1, 2),
                            [12.0 * nb_s[s] + 4.0 * nb_s[s] + 8.0 * fp[s] for s in range(S)])):
        ms = [evs[s][idx[0]].elapsed_time(evs[s][idx[1]]) for s in range(S)]
        avg_ms, avg_b = float(np.mean(ms)), float(np.mean(byt))
        ach = avg_b / (avg_ms * 1e-3) / 1e9
        kernels.append({"kernel": name, "avg_launch_ms": avg_ms, "alg_bytes_per_launch": avg_b,
                        "achieved": ach, "frac": ach / HBM_PEAK_GBS})
    log(args, "; ".join(f"{k_['kernel']}: {k_['avg_launch_ms']:.1f} ms, "
                        f"{k_['achieved']:.0f} GB/s alg" for k_ in kernels))
    dom = max(kernels, key=lambda k_: k_["avg_launch_ms"])
    del g, step
    plans = None

    if not args.no_plain:
        # per frame: the same planned kernels launched frame by frame
        fplans = [[ctx.plan(proj, im1, Lf[f][s], tile_shape=tune["tile_shape"])
                   for s in range(S)] for f in range(F)]

        def per_frame():
            for s in range(S):
                for f in range(F):
                    ctx.fill(u1[f], 0.0)
                    ctx.em_ratio(proj, im1, x1[f], Lf[f][s], vbuf, meas=per[f][s][1],
                                 plan=fplans[f][s])
                    ctx.back(proj, im1, u1[f], Lf[f][s], vbuf, plan=fplans[f][s])
                    ctx.em_update(x1[f], u1[f], sens, im1.nvox)

        res["per_frame_tiled_s"] = timed(per_frame, iters)
        rel["per_frame_tiled_vs_batched"] = rel_l2_t(xF, xb)
        log(args, f"DD per frame, tile plans: {res['per_frame_tiled_s']:.3f} s/iter, "
                  f"rel L2 vs batched {rel['per_frame_tiled_vs_batched']:.2e}")
        del fplans

        def plain():
            for s in range(S):
                for f in range(F):
                    ctx.fill(u1[f], 0.0)
                    ctx.em_subset(proj, im1, x1[f], u1[f], Lf[f][s], meas=per[f][s][1])
                    ctx.em_update(x1[f], u1[f], sens, im1.nvox)

        proj.set_accumulation(True)
        res["plain_s"] = timed(plain, iters)
        rel["plain_vs_batched"] = rel_l2_t(xF, xb)
        log(args, f"DD plain (per frame, fused EM): {res['plain_s']:.3f} s/iter, "
                  f"rel L2 vs batched {rel['plain_vs_batched']:.2e}")
    parity = None if args.no_oracle_sample else dd_oracle_sample(
        args, torch, ctx, proj, sc, im1, pp, per[0][0][0], per[0][0][1], xb[0:1], dev)
    # the CPU baseline samples the largest frame's first subset (~cpu_seconds)
    fbig = max(range(F), key=lambda f: per[f][0][0].numel())
    cpu = None if args.no_cpu_baseline else dd_cpu_baseline(
        args, torch, sc, im1, pp, per[fbig][0][0], per[fbig][0][1], xb[fbig:fbig + 1], fbig)
    # measured traffic applies to the profiled configuration only
    traffic, traffic_src = None, None
    if (args.dd_min == 1e6 and args.dd_max == 40e6 and S == 12 and args.order == "lor"):
        traffic = profiled_traffic(TRAFFIC_PROFILES["dd"], dom["kernel"])
        if traffic is not None:
            traffic_src = os.path.relpath(TRAFFIC_PROFILES["dd"], REPO)
    out = {
        "metric": "fwd+back event-projections/s, s/OSEM iter; batched DD+PSF histogram frames "
                  "vs plain per-frame",
        "value": 2.0 * n / res["batched_s"],
        "unit": "event-projections/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["batched_s"] * 1e3,
        "s_per_osem_iter": res["batched_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded dynamic phantom; per-frame Histogram3D of sampled counts, "
                "the nonzero bins projected)",
        "config": {"workload": f"DD + projection PSF (2 mm FWHM) histogram OSEM, 48x384 scanner "
                               f"R350, 128x128x96 @2.5mm, {F} frames log-uniform "
                               f"[{args.dd_min:g}, {args.dd_max:g}] counts, {S} subsets",
                   "frame_counts": counts, "nonzero_bins": n, "footprint_voxels": sum(fp)},
        "roofline": {"bound": "hbm", "achieved": dom["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["frac"], "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": dom["kernel"], "avg_launch_ms": dom["avg_launch_ms"],
                     "alg_bytes_per_launch": dom["alg_bytes_per_launch"],
                     "footprint_voxels_per_bin": sum(fp) / n},
        "kernels": kernels,
        "footprint_voxels_per_s": 2.0 * sum(fp) / res["batched_s"],
        "plan_bytes": plan_bytes, "plan_records": plan_records, "tuning": tune,
        "seconds": res, "rel_l2": rel, "parity": parity, "cpu_baseline": cpu,
    }
    if "plain_s" in res:
        out["plain_value"] = 2.0 * n / res["plain_s"]
        out["speedup_vs_plain"] = res["plain_s"] / res["batched_s"]
        out["speedup_vs_per_frame_tiled"] = res["per_frame_tiled_s"] / res["batched_s"]
    return out


def dd_bins_oracle(sc, bins, meas):
    """Oracle LORs of histogram bins (Histogram3D bin -> detector pair)."""
    import oracle as O
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_scanner
    osc = oracle_scanner(sc)
    pos, ori = O.lut(osc)
    d1, d2 = O.histo_pairs(osc, bins.cpu().numpy().astype(np.int64))
    return O.lors_from_pairs(pos, ori, d1, d2), meas.cpu().numpy()


def dd_oracle_sample(args, torch, ctx, proj, sc, im, pp, bins, meas, x, dev):
    """The EM update image and forward projections of an evenly strided
    sample of frame 0's first subset, GPU (tile-planned kernels) against the
    oracle's FP32 and FP64 restatements of ProjectorDD + the EM subset."""
    import oracle as O
    import yrtmi
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_image, oracle_proj, rel_l2
    k = min(args.oracle_events // 10, bins.numel())
    idx = strided_sample(torch, bins.numel(), k, dev)
    b, m = bins[idx].contiguous(), meas[idx].contiguous()
    L = yrtmi.LorSet.bins_(b)
    y = torch.empty(k, device=dev)
    ctx.forward(proj, im, x, L, y)
    v = torch.empty(k, device=dev)
    ug = torch.zeros(im.shape, device=dev)
    pl = ctx.plan(proj, im, L)
    ctx.em_ratio(proj, im, x, L, v, meas=m, plan=pl)
    ctx.back(proj, im, ug, L, v, plan=pl)
    ctx.synchronize()
    nthreads = max(1, min(16, os.cpu_count() or 1))
    OL, mh = dd_bins_oracle(sc, b, m)
    po, oi = oracle_proj(pp), oracle_image(im)
    xh = x.cpu().numpy().reshape(im.shape)
    t0 = time.perf_counter()
    y32 = O.forward(oi, xh, po, OL, nthreads=nthreads)
    y64 = O.forward(oi, xh.astype(np.float64), po, OL, dtype=np.float64, nthreads=nthreads)
    u32 = O.em_subset(oi, xh, np.zeros(im.shape, np.float32), po, OL, meas=mh,
                      nthreads=nthreads)
    u64 = O.em_subset(oi, xh.astype(np.float64), np.zeros(im.shape, np.float64), po, OL,
                      meas=mh, dtype=np.float64, nthreads=nthreads)
    dt = time.perf_counter() - t0
    yg, ugh = y.cpu().numpy(), ug.cpu().numpy()
    res = {"bins": k,
           "sample": f"{k} nonzero bins evenly strided over frame 0's first subset, the "
                     f"frame's image after 4 batched iterations",
           "fwd_rel_l2_gpu_vs_oracle_f32": rel_l2(yg, y32),
           "fwd_rel_l2_gpu_vs_oracle_f64": rel_l2(yg, y64),
           "fwd_rel_l2_oracle_f32_vs_f64": rel_l2(y32, y64),
           "em_rel_l2_gpu_vs_oracle_f32": rel_l2(ugh, u32),
           "em_rel_l2_gpu_vs_oracle_f64": rel_l2(ugh, u64),
           "em_rel_l2_oracle_f32_vs_f64": rel_l2(u32, u64),
           "oracle_s": dt}
    log(args, "DD oracle sample: " + json.dumps(res))
    return res


def dd_cpu_baseline(args, torch, sc, im, pp, bins, meas, x, frame=0):
    """The oracle's FP32 restatement of the reference's DD + PSF EM subset
    (std::thread, one per hardware thread up to 16) on an evenly strided
    sample of one frame's first subset, sized to ~cpu_seconds of work."""
    import oracle as O
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_image, oracle_proj
    nthreads = max(1, min(16, os.cpu_count() or 1))
    po, oi = oracle_proj(pp), oracle_image(im)
    xh = x.cpu().numpy().reshape(im.shape)

    def run(k):
        idx = strided_sample(torch, bins.numel(), k, bins.device)
        OL, mh = dd_bins_oracle(sc, bins[idx], meas[idx])
        uh = np.zeros(im.shape, np.float32)
        t0 = time.perf_counter()
        O.em_subset(oi, xh, uh, po, OL, meas=mh, nthreads=nthreads)
        return time.perf_counter() - t0

    k0 = 5000
    dt0 = run(k0)
    k = int(min(bins.numel(), max(k0, k0 / dt0 * args.cpu_seconds)))
    dt = run(k)
    log(args, f"DD cpu baseline: {k} bins in {dt:.1f}s on {nthreads} threads")
    return {"value": 2.0 * k / dt, "unit": "event-projections/s", "cores": nthreads,
            "kind": "port",
            "sample": f"DD + PSF EM subset (forward+backprojection) of {k} nonzero bins evenly "
                      f"strided over frame {frame}'s first subset, oracle FP32 restatement of the "
                      f"reference CPU path, {nthreads} std::threads",
            "seconds": dt, **host_cpu()}


def bench_autograd(args, torch, dev, stream):
    """The differentiable TOF list-mode projector behind torch.autograd
    (the 8-image configuration): a batch of 8 images 128x128x64 (2 mm
    voxels: the configuration leaves the size open), each a sum of 10
    Gaussian blobs (sigma U[3, 15] mm, amplitude U[0.5, 3]); 20 M events per
    image drawn from it (TOF 350 ps FWHM, 4 rays), the batch one launch (the
    image index as each event's frame).  One step: the Poisson negative
    log-likelihood sum(ybar + b - log(ybar + b)) over the events (list-mode
    counts of 1; b = randoms + scatter, a constant 10 % of the initial
    ybar's mean) of the image parameters, forward (A x) and backward
    (x.grad = A^T of the loss gradient, the tile-planned backprojection) in
    torch.autograd; then the dot-product (adjoint) check <A x, y> = <x, A^T y>
    on the whole batch in FP32."""
    import yrtmi
    from yrtmi import autograd as AG
    from yrtmi import synth

    t_setup = time.time()
    sc, _ = scanner_and_image()
    B, per = 8, args.ag_events
    im = yrtmi.ImageParams(128, 128, 64, 2.0, 2.0, 2.0, nt=B)
    im1 = yrtmi.ImageParams(128, 128, 64, 2.0, 2.0, 2.0)
    rng = np.random.default_rng(args.seed + 4)
    imgs = [synth.gaussian_blobs(im1, rng) for _ in range(B)]
    evs = [synth.sample_events(sc, im1, imgs[i], per, seed=args.seed + 17 * i,
                               tof_fwhm_ps=350.0, device=dev) for i in range(B)]
    L = AG.batch_lors([e["d1"] for e in evs], [e["d2"] for e in evs], [e["tof"] for e in evs])
    del evs
    ctx = yrtmi.Context(sc, stream=stream)
    pp = yrtmi.ProjectorParams(sc, numRays=4)
    pp.addTOF(350.0, 5)
    proj = ctx.projector(pp)
    n = L.n
    x0 = torch.ones(im.shape, device=dev)
    if args.order == "lor":
        L = L.permuted(ctx, ctx.lor_order(im, L))  # frame-major locality order
    tune = ctx.autotune(proj, im, L, x0)
    t0 = time.time()
    plan = ctx.plan(proj, im, L, tile_shape=tune["tile_shape"], persistent=tune["persistent"])
    ctx.synchronize()
    plan_s = time.time() - t0
    op = AG.ProjectionOperator(ctx, proj, im, L, plan=plan)
    segs = ctx.count_segments(proj, im, L)
    log(args, f"autograd: {B} images, {n} events, {segs / n:.1f} segments/event, plan "
              f"{plan.info()['bytes'] / 2**30:.1f} GiB in {plan_s:.2f}s; setup "
              f"{time.time() - t_setup:.1f}s; {tune}")
    x = torch.nn.Parameter(x0.clone())
    with torch.no_grad():
        b = 0.1 * float(op.forward(x0).mean())

    def step(evs=None):
        x.grad = None
        if evs is not None:
            evs[0].record()
        ybar = AG.project(x, op)
        if evs is not None:
            evs[1].record()
        loss = AG.poisson_nll(ybar, 1.0, b)
        loss.backward()
        if evs is not None:
            evs[2].record()
        return loss

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        loss = step()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    evs = tuple(torch.cuda.Event(enable_timing=True) for _ in range(3))
    step(evs)
    torch.cuda.synchronize()
    fwd_ms, back_ms = evs[0].elapsed_time(evs[1]), evs[1].elapsed_time(evs[2])
    # dot-product test on the batch: <A u, w> vs <u, A^T w>, FP64 sums
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    u = torch.rand(im.shape, device=dev, generator=g)
    w = torch.rand(n, device=dev, generator=g)
    lhs = float((op.forward(u).double() * w.double()).sum())
    rhs = float((u.double() * op.adjoint(w).double()).sum())
    adj = abs(lhs - rhs) / abs(lhs)
    log(args, f"{wall * 1e3:.1f} ms/step (forward {fwd_ms:.1f} ms, backward {back_ms:.1f} ms); "
              f"adjoint rel diff {adj:.2e}; loss {float(loss.detach()):.6e}")
    # algorithmic bytes: forward 12 B/event + 4 B per segment + 4 B stored;
    # backward 12 + 4 B/event + 8 B per segment
    kern = [("forward A x (kEvents<MODE_FWD>)", fwd_ms, 16.0 * n + 4.0 * segs),
            (f"backward A^T (kTileBack<TOF,shape={tune['tile_shape']}> + loss gradient)",
             back_ms, 16.0 * n + 8.0 * segs)]
    kernels = [{"kernel": k_, "avg_launch_ms": ms, "alg_bytes_per_launch": by,
                "achieved": by / (ms * 1e-3) / 1e9, "frac": by / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
               for k_, ms, by in kern]
    dom = max(kernels, key=lambda k_: k_["avg_launch_ms"])
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, sc, im1, pp, L.d1, L.d2, L.tof, [0, n], x0[0:1])
    return {
        "metric": "fwd+back event-projections/s (autograd step of a Poisson NLL)",
        "value": 2.0 * n / wall,
        "unit": "event-projections/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian-blob phantoms, physically sampled TOF events)",
        "config": {"workload": f"differentiable TOF list-mode projector (autograd), batch of {B} "
                               f"images 128x128x64 @2mm, {per / 1e6:g}M events each, 4-ray "
                               f"Siddon TOF 350 ps, Poisson NLL with 10% background",
                   "events": n, "segments_per_event": segs / n, "tuning": tune},
        "roofline": {"bound": "hbm", "achieved": dom["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["frac"], "traffic": None,
                     "kernel": dom["kernel"], "avg_launch_ms": dom["avg_launch_ms"],
                     "alg_bytes_per_launch": dom["alg_bytes_per_launch"]},
        "kernels": kernels,
        "adjoint_rel_diff_fp32": adj,
        "plan_s": plan_s,
        "cpu_baseline": cpu,
    }


def bench_histogram(args, torch, dev, stream):
    """Histogram OSEM, full size, GPU (C++ host layer) and CPU oracle."""
    import oracle as O
    import pyyrtmi as Y
    import yrtmi
    from yrtmi import synth
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_image, oracle_scanner, rel_l2

    a = ("cfg1", 96.0, 4.0, 4.0, 20.0, 300.0, 256, 24)
    sc, ysc = yrtmi.Scanner(*a, maxRingDiff=23), Y.Scanner(*a, 1, 23, 1, 1)
    im = yrtmi.ImageParams(96, 96, 48, 2.0, 2.0, 2.0)
    yim = Y.ImageParams.fromParams(96, 96, 48, 2.0, 2.0, 2.0)
    S, iters = 8, 5
    rng = np.random.default_rng(args.seed)
    spheres = synth.random_spheres_in_cylinder(rng, 100.0, 90.0, [5.0, 8.0, 11.0, 15.0], 4.0)
    act = synth.cylinder_spheres(im, 100.0, 90.0, 1.0, spheres, supersample=2)
    body = synth.cylinder_spheres(im, 100.0, 90.0, 1.0, [], supersample=1)
    mu = np.where(body > 0, np.float32(0.096), np.float32(0.0)).astype(np.float32)  # cm^-1
    # measured histogram: Poisson(attenuated forward projection, 50 M counts)
    ctx = yrtmi.Context(sc, stream=stream)
    pp = yrtmi.ProjectorParams(sc)
    proj = ctx.projector(pp)
    L = yrtmi.LorSet.histo_subset(ctx, 1, 0)
    act_t = torch.from_numpy(act.astype(np.float32)).to(dev)
    mu_t = torch.from_numpy(mu).to(dev)
    y = torch.empty(L.n, device=dev)
    ctx.forward(proj, im, act_t, L, y)
    acf = torch.empty(L.n, device=dev)
    ctx.attenuation_factors(im, mu_t, L, acf)
    y = y * acf
    y = y * (50e6 / float(y.sum()))
    g = torch.Generator(device=dev)
    g.manual_seed(args.seed)
    y = torch.poisson(y, generator=g).contiguous()
    ctx.synchronize()
    counts = float(y.sum())
    log(args, f"histogram: {L.n} bins, {counts:.3e} counts, {int((y > 0).sum())} nonzero")
    # GPU: the C++ host layer's OSEM
    osem = Y.createOSEM(ysc)
    yctx = osem.getContext()
    h = Y.Histogram3DDeviceAlias(yctx, y)
    mu_img = Y.ImageDeviceAlias(yctx, yim, mu_t)
    osem.setImageParams(yim)
    osem.setAttenuationImage(mu_img)
    osem.num_OSEM_subsets = S
    osem.num_MLEM_iterations = iters
    osem.setDataInput(h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    osem.generateSensitivityImages()
    yctx.synchronize()
    t_sens = time.perf_counter() - t0
    t0 = time.perf_counter()
    x = osem.reconstruct()
    t_rec = time.perf_counter() - t0
    xg = x.transferToHostMemory()
    log(args, f"GPU: sensitivity {t_sens:.3f} s, reconstruct ({iters} x {S}) {t_rec:.3f} s")
    # CPU: the oracle's restatement of the reference CPU OSEM on the same data
    nthreads = max(1, min(16, os.cpu_count() or 1))
    osc = oracle_scanner(sc)
    oi = oracle_image(im)
    po = O.proj_for_scanner(osc)
    pos, ori = O.lut(osc)
    zb, ph, nr = O.histo_dims(osc)
    assert zb * ph * nr == L.n, "histogram sizes differ"
    d1, d2 = O.histo_pairs(osc, np.arange(L.n))
    OL = O.lors_from_pairs(pos, ori, d1, d2)
    yh = y.cpu().numpy()
    if args.no_cpu_baseline:
        return _histo_result(L.n, counts, iters, S, t_rec, t_sens, None)
    t0 = time.perf_counter()
    osens = O.sensitivity_images(osc, oi, po, S, False, att_im=oi, att_img=mu, nthreads=nthreads,
                                 pos_ori=(pos, ori))
    c_sens = time.perf_counter() - t0
    acf_o = O.attenuation_factors(oi, mu, OL, nthreads)
    subsets = [O.histo_subset_bins(osc, S, s) for s in range(S)]
    t0 = time.perf_counter()
    xo = O.reconstruct(osc, oi, po, OL, osens, iters, S, False, meas=yh, att=acf_o,
                       subset_indices=subsets, nthreads=nthreads)
    c_rec = time.perf_counter() - t0
    r = rel_l2(xg, xo)
    log(args, f"CPU ({nthreads} threads): sensitivity {c_sens:.1f} s, reconstruct {c_rec:.1f} s; "
              f"final images rel L2 {r:.2e}")
    cpu = {"value": L.n * iters / c_rec, "unit": "bins/s", "cores": nthreads, "kind": "port",
           "sample": "the full reconstruction (5 iterations x 8 subsets) by the oracle's FP32 "
                     "restatement of the reference CPU OSEM",
           "seconds": c_rec, "s_per_osem_iter": c_rec / iters, "sensitivity_s": c_sens,
           **host_cpu()}
    out = _histo_result(L.n, counts, iters, S, t_rec, t_sens, cpu)
    out["parity_rel_l2_vs_cpu"] = r
    return out


def _histo_result(nbins, counts, iters, S, t_rec, t_sens, cpu):
    return {
        "metric": "histogram bins reconstructed/s, s/OSEM iter (histogram OSEM)",
        "value": nbins * iters / t_rec,
        "unit": "bins/s",
        "n_gpus": 1, "steps": iters, "warmup": 0,
        "ms_per_step": t_rec / iters * 1e3,
        "s_per_osem_iter": t_rec / iters,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded phantom; Poisson histogram of 50 M counts)",
        "config": {"workload": "histogram OSEM, Siddon, no TOF, 24x256 scanner R300, "
                               f"96x96x48 @2mm, {iters} iterations x {S} subsets, attenuation",
                   "bins": nbins, "counts": counts,
                   "bins_counted": "every histogram bin of an iteration; the GPU path skips "
                                   "zero-count bins, the CPU path projects them"},
        "sensitivity_s": t_sens,
        "cpu_baseline": cpu,
    }


def host_cpu():
    """nproc and the CPU model of the machine the CPU baseline runs on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(args, sc, im, pp, d1, d2, tof, bounds, x, mot=None, xf=None):
    """The CPU restatement of the reference EM subset (oracle, std::thread
    over contiguous event chunks like util::parallelForChunked) on an evenly
    strided sample of subset 0 (every motion frame represented), sized to
    ~cpu_seconds of work."""
    import oracle as O
    import torch
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_image, oracle_proj, oracle_scanner
    nthreads = max(1, min(16, os.cpu_count() or 1))
    osc = oracle_scanner(sc)
    pos, ori = O.lut(osc)
    po = oracle_proj(pp)
    oi = oracle_image(im)
    xh = x.cpu().numpy()

    def run(k):
        idx = strided_sample(torch, bounds[1], k, d1.device)
        a, b, t = (d1[idx].cpu().numpy().astype(np.uint32),
                   d2[idx].cpu().numpy().astype(np.uint32), tof[idx].cpu().numpy())
        m = None if mot is None else mot[idx].cpu().numpy()
        L = O.lors_from_pairs(pos, ori, a, b, tof=t, motion_frame=m, motion_xf=xf)
        uh = np.zeros(im.shape, np.float32)
        t0 = time.perf_counter()
        O.em_subset(oi, xh, uh, po, L, nthreads=nthreads)
        return time.perf_counter() - t0

    k0 = 20000
    dt0 = run(k0)
    k = int(min(bounds[1], max(k0, k0 / dt0 * args.cpu_seconds)))
    dt = run(k)
    rate = 2.0 * k / dt
    log(args, f"cpu baseline: {k} events in {dt:.1f}s on {nthreads} threads")
    return {"value": rate, "unit": "event-projections/s", "cores": nthreads, "kind": "port",
            "sample": f"EM subset (forward+backprojection, LOR motion applied) of {k} events "
                      f"evenly strided over subset 0, oracle FP32 restatement of the reference "
                      f"CPU path, {nthreads} std::threads" if mot is not None else
                      f"EM subset (forward+backprojection) of {k} events evenly strided over "
                      f"subset 0, oracle FP32 restatement, {nthreads} std::threads",
            "seconds": dt, **host_cpu()}


if __name__ == "__main__":
    sys.exit(main())
