#!/usr/bin/env python3
"""Benchmark: list-mode TOF OSEM with 4-ray Siddon on one MI355X.

Default workload (the 64-ring x 512-crystal scanner configuration): scanner
radius 400 mm, 3.2 x 3.2 x 20 mm crystals, 100 M list-mode events (uint32 d1,
d2 + float32 TOF) drawn from a seeded phantom (water cylinder r = 150 mm,
length 180 mm, plus 6 spheres of radius 4-14 mm at 5:1 contrast), TOF 250 ps
FWHM truncated at 5 sigma, image 200 x 200 x 128 at 2 mm, 10 subsets.

One step = one full OSEM iteration: for each of the 10 subsets, the EM
update image of that subset's events (forward projection and ratio per
event, then backprojection) followed by the multiplicative image update.
Default (--accum tiled): the backprojection goes through per-subset tile
plans (the Siddon walks binned by image tiles, accumulated in LDS), built
once in setup, with the launch shapes picked by the autotuner on subset 0
(--no-tune: defaults), and the iteration replayed as one recorded graph
(--no-graph: launched call by call).  --accum xcd / direct run the fused
one-kernel EM with global atomics instead.  Events, plans, the detector LUT
and all images are resident in HBM before timing starts; the sensitivity
image (backprojection of all 536.9 M scanner LORs x 4 rays), the plans and
the tuning are made once in setup and timed separately.  Per-kernel times
(roofline) come from one more iteration run call by call with HIP events on
the context's stream.

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

--workload motion: list-mode TOF OSEM with LOR motion: 250M events in 8
rigid motion frames (contiguous stretches of the acquisition), 256x256x128
@1.6 mm, 10 subsets, 4 rays; the sensitivity image time-averaged over the
motion.  No CPU baseline (the default workload carries it).

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
# Per-launch HBM traffic of the default list-mode configuration, from two
# rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) scaled by a copy of known
# size; the file gives the commands.
TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "em_listmode_tof_100M_pmc_traffic_sq.md")


def profiled_traffic(kernel_label):
    """HBM bytes per launch of `kernel_label` (its part before the first
    space) from the committed counter summary, or None."""
    import re
    try:
        text = open(TRAFFIC_PROFILE).read()
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
    ap.add_argument("--workload", choices=("listmode", "frames", "histogram", "dd", "motion",
                                           "batch"),
                    default="listmode")
    ap.add_argument("--dd-min", type=float, default=1e6, help="dd workload: min counts/frame")
    ap.add_argument("--dd-max", type=float, default=40e6, help="dd workload: max counts/frame")
    ap.add_argument("--dd-subsets", type=int, default=12)
    ap.add_argument("--events", type=int, default=100_000_000)
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
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


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
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not built)"}))
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
        else:
            out = bench_listmode(args, torch, dev, stream)
    print(json.dumps(out))
    return 0


def bench_listmode(args, torch, dev, stream):
    import yrtmi
    from yrtmi import synth

    t_setup = time.time()
    motion = args.workload == "motion"
    sc, im = scanner_and_image(motion)
    act = phantom(args, im)
    log(args, f"phantom ready ({time.time() - t_setup:.1f}s)")
    # the motion workload: 250M events in 8 rigid motion frames (contiguous
    # stretches of the acquisition) unless --events says otherwise
    n_events = 250_000_000 if motion and args.events == 100_000_000 else args.events
    NMF = 8
    ev = synth.sample_events(sc, im, act, n_events, seed=args.seed,
                             tof_fwhm_ps=250.0, device=dev)
    S, n = args.subsets, n_events
    if motion:
        ev["motion"] = (torch.arange(n, device=dev, dtype=torch.int64) * NMF // n).to(torch.uint8)
    ev, bounds = subset_major(torch, ev, n, S, dev)
    d1, d2, tof = ev["d1"], ev["d2"], ev["tof"]
    mot = ev.get("motion")
    del ev
    torch.cuda.synchronize()
    log(args, f"{n} events drawn ({time.time() - t_setup:.1f}s)")

    ctx = yrtmi.Context(sc, stream=stream)
    if motion:
        ctx.upload_motion(synth.rigid_transforms(np.random.default_rng(args.seed), NMF))
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
            pl = ctx.plan(psens, im, Ls, args.chunk, tile_shape=yrtmi.capi.TILE_32x32x8)
            ctx.sensitivity(psens, im, sens, Ls, plan=pl)
            del pl
        else:
            ctx.sensitivity(psens, im, sens, Ls)
    if motion:
        # the sensitivity image time-averaged over the motion frames (equal
        # durations: weight 1/8 each), as the reconstruct command does
        moved = torch.zeros_like(sens)
        ctx.time_average_move(im, sens, moved,
                              synth.rigid_transforms(np.random.default_rng(args.seed), NMF),
                              np.full(NMF, 1.0 / NMF, np.float32))
        sens = moved
    ctx.synchronize()
    sens_s = time.time() - t0
    n_lors = sum(yrtmi.LorSet.histo_subset(ctx, SS, s).n for s in range(SS))
    log(args, f"sensitivity: {n_lors} LORs x 4 rays in {sens_s:.2f}s")
    # hard threshold (voxels <= 1 -> 0), list-mode scaling 1/S, mask, init
    ctx.threshold(sens, sens.clone(), im.nvox, 1.0, 0.0, 0.0, 1.0, 0.0)
    sens *= float(np.float32(1.0) / np.float32(S))
    x = torch.full(im.shape, 0.125, device=dev)
    ctx.threshold(x, sens, im.nvox, 0.0, 0.0, 0.0, 1.0, 0.0)
    u = torch.zeros(im.shape, device=dev)
    sub = [yrtmi.LorSet.pairs(d1[bounds[s]:bounds[s + 1]], d2[bounds[s]:bounds[s + 1]],
                              tof=tof[bounds[s]:bounds[s + 1]],
                              motion_frame=None if mot is None else mot[bounds[s]:bounds[s + 1]])
           for s in range(S)]
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
    vbuf = None
    if tiled:
        shape, pers = args.tile_shape, 0
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
    # context's stream
    evs = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(S)]
    iteration(evs)
    ctx.synchronize()
    s_per_iter = wall / args.steps
    ev_proj_per_s = 2.0 * n * args.steps / wall
    # Algorithmic bytes per launch (R_ev = 12 B/event: d1, d2, tof; an atomic
    # counts as an 8 B read-modify-write):
    #   forward + ratio:  R_ev + 4 B per voxel segment + 4 B ratio stored
    #   backprojection:   R_ev + 4 B value + 8 B per voxel segment
    #   fused EM:         R_ev + 12 B per voxel segment
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
             [12.0 * nev[s] + 4.0 * segs[s] + 4.0 * nev[s] for s in range(S)])
        shape = tune["tile_shape"] if tune else args.tile_shape
        roof(f"kTileBack<TOF,shape={shape}> (tile-binned backprojection)", (1, 2),
             [12.0 * nev[s] + 4.0 * nev[s] + 8.0 * segs[s] for s in range(S)])
    else:
        roof(f"kEvents<MODE_EM,Siddon,TOF,LPE={lpe}> (fused EM)", (0, 2),
             [12.0 * nev[s] + 12.0 * segs[s] for s in range(S)])
    dom = max(kernels, key=lambda k: k["avg_launch_ms"])
    # measured traffic applies to the profiled configuration only
    traffic, traffic_src = None, None
    if (tiled and n == 100_000_000 and S == 10 and args.order == "lor"
            and not args.no_tune and not motion):
        traffic = profiled_traffic(dom["kernel"])
        if traffic is not None:
            traffic_src = os.path.relpath(TRAFFIC_PROFILE, REPO)
    log(args, f"{s_per_iter:.3f} s/iter ({'graph' if graph else 'calls'}); " + "; ".join(
        f"{k['kernel']}: {k['avg_launch_ms']:.2f} ms, {k['achieved']:.0f} GB/s alg" for k in kernels))

    cpu = None
    if not args.no_cpu_baseline and not motion:
        cpu = cpu_baseline(args, sc, im, pp, d1, d2, tof, bounds, x)
    return {
        "metric": "fwd+back event-projections/s, s/OSEM iter",
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
        "data": "synthetic (seeded phantom, physically sampled TOF list-mode events)",
        "config": {"workload": (f"list-mode TOF OSEM with LOR motion ({NMF} rigid frames), "
                                f"4-ray Siddon, 64x512 scanner R400, {n / 1e6:g}M events, "
                                f"256x256x128 @1.6mm, {S} subsets") if motion else
                               ("list-mode TOF OSEM, 4-ray Siddon, 64x512 scanner R400, "
                                "100M events, 200x200x128 @2mm, 10 subsets"),
                   "motion_frames": NMF if motion else 0,
                   "events": n, "subsets": S, "rays": 4, "tof_fwhm_ps": 250.0,
                   "tof_n_std": 5, "image": [im.nx, im.ny, im.nz], "voxel_mm": float(im.vx),
                   "lanes_per_event": lpe, "accumulation": args.accum,
                   "event_order": args.order,
                   "graph": graph is not None, "parallelism": "single GPU"},
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
        "cpu_baseline": cpu,
    }


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
    """Batched distance-driven OSEM with projection-space PSF over dynamic
    frames (the 48-ring configuration): plain per-frame vs one launch over all
    frames, direct and with tile plans (the value: batched with plans)."""
    import yrtmi
    from yrtmi import synth

    # 48 rings x 384 crystals, R 350 mm, 4 x 4 x 20 mm crystals
    sc = yrtmi.Scanner("cfg3", 192.0, 4.0, 4.0, 20.0, 350.0, 384, 48)
    im1 = yrtmi.ImageParams(128, 128, 96, 2.5, 2.5, 2.5)
    F, S = 16, args.dd_subsets
    rng = np.random.default_rng(args.seed + 3)
    counts = [int(c) for c in np.exp(rng.uniform(math.log(args.dd_min), math.log(args.dd_max), F))]
    # activity: water cylinder (1.0) and 5 spheres whose uptake follows a
    # one-tissue compartment curve 4 (1 - exp(-k2 t)) / (1 - exp(-k2 T)),
    # k2 ~ U[0.01, 0.3] / s, frames at t = 60 s (f + 1)
    spheres = synth.random_spheres_in_cylinder(rng, 120.0, 160.0, [6.0, 9.0, 12.0, 15.0, 18.0], 4.0)
    k2 = rng.uniform(0.01, 0.3, len(spheres))
    T = 60.0 * F
    t0 = time.time()
    evs = []
    for f in range(F):
        t = 60.0 * (f + 1)
        sph = [(cx, cy, cz, r, 1.0 + 3.0 * (1 - math.exp(-k * t)) / (1 - math.exp(-k * T)))
               for (cx, cy, cz, r, _), k in zip(spheres, k2)]
        act = synth.cylinder_spheres(im1, 120.0, 160.0, 1.0, sph, supersample=1)
        evs.append(synth.sample_events(sc, im1, act, counts[f], seed=args.seed + 11 * f,
                                       device=dev))
    n = sum(counts)
    log(args, f"DD: {F} frames, {n} events ({min(counts)}..{max(counts)}) in {time.time() - t0:.1f}s")
    ctx = yrtmi.Context(sc, stream=stream)
    pp = yrtmi.ProjectorParams(sc, projectorType="DD")
    k, s_step, ks = synth.gaussian_psf_row(2.0, k_spacing=0.25, num_s=1, s_step=1000.0)
    pp.psf_kernels, pp.psf_s_step, pp.psf_k_spacing = k, s_step, ks
    proj = ctx.projector(pp)
    proj.set_accumulation(True)
    sens = torch.full(im1.shape, 1.0, device=dev)
    per = []
    for f in range(F):
        e, b = subset_major(torch, evs[f], counts[f], S, dev)
        per.append([yrtmi.LorSet.pairs(e["d1"][b[s]:b[s + 1]], e["d2"][b[s]:b[s + 1]])
                    for s in range(S)])
    del evs
    imF = yrtmi.ImageParams(im1.nx, im1.ny, im1.nz, im1.vx, im1.vy, im1.vz, nt=F)
    xF = torch.full(imF.shape, 0.125, device=dev)
    uF = torch.zeros(imF.shape, device=dev)
    x1 = [xF[f:f + 1] for f in range(F)]
    u1 = [uF[f:f + 1] for f in range(F)]
    # footprint voxels (the DD work unit), exactly
    fp = sum(ctx.count_segments(proj, im1, per[f][s]) for f in range(F) for s in range(S))

    def plain():
        for f in range(F):
            for s in range(S):
                ctx.fill(u1[f], 0.0)
                ctx.em_subset(proj, im1, x1[f], u1[f], per[f][s])
                ctx.em_update(x1[f], u1[f], sens, im1.nvox)

    cat = lambda key, s: torch.cat([getattr(per[f][s], key) for f in range(F)]).contiguous()
    bsub = []
    for s in range(S):
        fr = torch.cat([torch.full((per[f][s].n,), f, dtype=torch.int32, device=dev)
                        for f in range(F)])
        Lb = yrtmi.LorSet.pairs(cat("d1", s), cat("d2", s), frame=fr)
        if args.order == "lor":
            Lb = Lb.permuted(ctx, ctx.lor_order(imF, Lb))
        bsub.append(Lb)

    def batched():
        for s in range(S):
            ctx.fill(uF, 0.0)
            ctx.em_subset(proj, imF, xF, uF, bsub[s])
            ctx.em_update(xF, uF, sens, im1.nvox)

    def timed(fn, steps):
        fn()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        ctx.synchronize()
        return (time.perf_counter() - t) / steps

    res = {"plain_s": timed(plain, args.steps)}
    log(args, f"DD plain (per frame, fused EM): {res['plain_s']:.3f} s/iter")
    batched()
    ctx.synchronize()
    with ctx.graph() as g:
        batched()
    res["batched_direct_s"] = timed(g.launch, args.steps)
    log(args, f"DD batched (frames in one launch, locality order, graph): "
              f"{res['batched_direct_s']:.3f} s/iter")
    del g
    # tile plans of the batched subsets (all frames' tiles: the 32x32x8 shape
    # keeps 16 frames within the bin limit)
    ctx.synchronize()
    t0 = time.perf_counter()
    plans = [ctx.plan(proj, imF, bsub[s], tile_shape=yrtmi.capi.TILE_32x32x8) for s in range(S)]
    ctx.synchronize()
    res["plan_s"] = time.perf_counter() - t0
    plan_bytes = sum(p_.info()["bytes"] for p_ in plans)
    plan_records = sum(p_.info()["records"] for p_ in plans)

    def batched_planned():
        for s in range(S):
            ctx.fill(uF, 0.0)
            ctx.em_subset(proj, imF, xF, uF, bsub[s], plan=plans[s])
            ctx.em_update(xF, uF, sens, im1.nvox)

    batched_planned()
    ctx.synchronize()
    with ctx.graph() as g:
        batched_planned()
    res["batched_s"] = timed(g.launch, args.steps)
    log(args, f"DD batched + tile plans ({plan_records / n:.1f} records/LOR, "
              f"{plan_bytes / 2**30:.2f} GiB, built in {res['plan_s']:.2f}s): "
              f"{res['batched_s']:.3f} s/iter")
    return {
        "metric": "fwd+back event-projections/s, batched DD+PSF dynamic frames vs plain per-frame",
        "value": 2.0 * n / res["batched_s"],
        "unit": "event-projections/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": res["batched_s"] * 1e3,
        "s_per_osem_iter": res["batched_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded dynamic phantom; each count one LOR)",
        "config": {"workload": f"DD + projection PSF (2 mm FWHM), 48x384 scanner R350, "
                               f"128x128x96 @2.5mm, {F} frames log-uniform "
                               f"[{args.dd_min:g}, {args.dd_max:g}] counts, {S} subsets",
                   "frame_counts": counts, "events": n, "footprint_voxels": fp},
        "plain_value": 2.0 * n / res["plain_s"],
        "speedup_vs_plain": res["plain_s"] / res["batched_s"],
        "footprint_voxels_per_s": 2.0 * fp / res["batched_s"],
        "plan_bytes": plan_bytes, "plan_records": plan_records,
        "seconds": res,
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


def cpu_baseline(args, sc, im, pp, d1, d2, tof, bounds, x):
    """The CPU restatement of the reference EM subset (oracle, std::thread
    over contiguous event chunks like util::parallelForChunked) on the first
    events of subset 0, sized to ~cpu_seconds of work."""
    import oracle as O
    sys.path.insert(0, os.path.join(REPO, "tests"))
    from helpers import oracle_image, oracle_proj, oracle_scanner
    nthreads = max(1, min(16, os.cpu_count() or 1))
    osc = oracle_scanner(sc)
    pos, ori = O.lut(osc)
    po = oracle_proj(pp)
    oi = oracle_image(im)
    xh = x.cpu().numpy()

    def run(k):
        a, b, t = (d1[:k].cpu().numpy().astype(np.uint32), d2[:k].cpu().numpy().astype(np.uint32),
                   tof[:k].cpu().numpy())
        L = O.lors_from_pairs(pos, ori, a, b, tof=t)
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
            "sample": f"EM subset (forward+backprojection) of the first {k} events of subset 0, "
                      f"oracle FP32 restatement, {nthreads} std::threads",
            "seconds": dt, **host_cpu()}


if __name__ == "__main__":
    sys.exit(main())
