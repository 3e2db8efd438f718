#!/usr/bin/env python3
"""Benchmark: list-mode TOF OSEM with 4-ray Siddon on one MI355X.

Workload (the 64-ring x 512-crystal scanner configuration): scanner radius
400 mm, 3.2 x 3.2 x 20 mm crystals, 100 M list-mode events (uint32 d1, d2 +
float32 TOF) drawn from a seeded phantom (water cylinder r = 150 mm, length
180 mm, plus 6 spheres of radius 4-14 mm at 5:1 contrast), TOF 250 ps FWHM
truncated at 5 sigma, image 200 x 200 x 128 at 2 mm, 10 subsets.

One step = one full OSEM iteration: for each of the 10 subsets, the EM
update image of that subset's events (forward projection and ratio per
event, then backprojection) followed by the multiplicative image update.
Default (--accum tiled): the backprojection goes through per-subset tile
plans (the Siddon walks binned by 32x32x16 image tiles, accumulated in LDS),
built once in setup; --accum xcd / direct run the fused one-kernel EM with
global atomics instead.  Events, plans, the detector LUT and all images are
resident in HBM before timing starts; the sensitivity image
(backprojection of all 536.9 M scanner LORs x 4 rays) and the plans are
made once in setup and timed separately.

Prints one JSON line (see the `out` dict at the end).
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=int, default=100_000_000)
    ap.add_argument("--subsets", type=int, default=10)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--lpe", type=int, default=4, help="lanes per event")
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--accum", choices=("tiled", "xcd", "direct"), default="tiled",
                    help="backprojection: tile plans with LDS accumulation, or global "
                         "atomics into per-XCD images / the target directly")
    ap.add_argument("--chunk", type=int, default=0, help="tile plan records per workgroup")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 256 MiB device copy first (PMC calibration)")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def log(args, *a):
    if not args.quiet:
        print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: --gpus must be 1 "
                                   "(multi-GPU sharding is described in DESIGN.md, not built)"}))
        return 2
    import torch
    import yrtmi
    from yrtmi import synth

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calib_copy:
        # a plain copy of known size: 256 MiB read + 256 MiB written
        a = torch.ones(64 << 20, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    t_setup = time.time()
    sc = yrtmi.Scanner("cfg2", 64 * 3.2, 3.2, 3.2, 20.0, 400.0, 512, 64)
    im = yrtmi.ImageParams(200, 200, 128, 2.0, 2.0, 2.0)
    rng = np.random.default_rng(args.seed)
    spheres = synth.random_spheres_in_cylinder(rng, 150.0, 180.0,
                                               list(np.linspace(4.0, 14.0, 6)), 5.0)
    act = synth.cylinder_spheres(im, 150.0, 180.0, 1.0, spheres, supersample=2)
    log(args, f"phantom ready ({time.time() - t_setup:.1f}s)")
    ev = synth.sample_events(sc, im, act, args.events, seed=args.seed,
                             tof_fwhm_ps=250.0, device=dev)
    torch.cuda.synchronize()
    log(args, f"{args.events} events drawn ({time.time() - t_setup:.1f}s)")
    # subset-major permutation: event i belongs to subset i mod S
    # (BinIteratorChronologicalInterleaved); each subset becomes contiguous
    S = args.subsets
    n = args.events
    perm = torch.cat([torch.arange(s, n, S, device=dev) for s in range(S)])
    d1 = ev["d1"][perm].contiguous()
    d2 = ev["d2"][perm].contiguous()
    tof = ev["tof"][perm].contiguous()
    del ev, perm
    bounds = [0]
    for s in range(S):
        bounds.append(bounds[-1] + len(range(s, n, S)))

    ctx = yrtmi.Context(sc)
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
            pl = ctx.plan(psens, im, Ls, args.chunk)
            ctx.sensitivity(psens, im, sens, Ls, plan=pl)
            del pl
        else:
            ctx.sensitivity(psens, im, sens, Ls)
    torch.cuda.synchronize()
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
                              tof=tof[bounds[s]:bounds[s + 1]]) for s in range(S)]
    # algorithmic work: voxel segments per subset (instrumented count)
    segs = [ctx.count_segments(proj, im, L) for L in sub]
    log(args, f"segments/event {sum(segs) / n:.1f}; setup {time.time() - t_setup:.1f}s")
    plans, plan_s, plan_bytes, plan_records = [], 0.0, 0, 0
    vbuf = None
    if tiled:
        torch.cuda.synchronize()
        t0 = time.time()
        plans = [ctx.plan(proj, im, L, args.chunk) for L in sub]
        torch.cuda.synchronize()
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
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        evs = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3))
               for _ in range(S)]
        iteration(evs)
        kernel_ms.append(evs)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
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
        ms = [evs[s][idx[0]].elapsed_time(evs[s][idx[1]]) for evs in kernel_ms for s in range(S)]
        by = [bytes_per_subset[s] for _ in kernel_ms for s in range(S)]
        avg_ms, avg_b = float(np.mean(ms)), float(np.mean(by))
        ach = avg_b / (avg_ms * 1e-3) / 1e9
        kernels.append({"kernel": name, "avg_launch_ms": avg_ms, "alg_bytes_per_launch": avg_b,
                        "achieved": ach, "frac": ach / HBM_PEAK_GBS})

    if tiled:
        roof("kEmRatio<TOF,LPE=4> (forward + ratio)", (0, 1),
             [12.0 * nev[s] + 4.0 * segs[s] + 4.0 * nev[s] for s in range(S)])
        roof("kTileBack<TOF> (tile-binned backprojection)", (1, 2),
             [12.0 * nev[s] + 4.0 * nev[s] + 8.0 * segs[s] for s in range(S)])
    else:
        roof("kEvents<MODE_EM,Siddon,TOF,LPE=4> (fused EM)", (0, 2),
             [12.0 * nev[s] + 12.0 * segs[s] for s in range(S)])
    dom = max(kernels, key=lambda k: k["avg_launch_ms"])
    log(args, f"{s_per_iter:.3f} s/iter; " + "; ".join(
        f"{k['kernel']}: {k['avg_launch_ms']:.2f} ms, {k['achieved']:.0f} GB/s alg" for k in kernels))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, sc, im, pp, d1, d2, tof, bounds, x)
    out = {
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
        "config": {"workload": "list-mode TOF OSEM, 4-ray Siddon, 64x512 scanner R400, "
                               "100M events, 200x200x128 @2mm, 10 subsets",
                   "events": n, "subsets": S, "rays": 4, "tof_fwhm_ps": 250.0,
                   "tof_n_std": 5, "image": [200, 200, 128], "voxel_mm": 2.0,
                   "lanes_per_event": args.lpe, "block": args.block,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": dom["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["frac"], "traffic": None,
                     "kernel": dom["kernel"], "avg_launch_ms": dom["avg_launch_ms"],
                     "alg_bytes_per_launch": dom["alg_bytes_per_launch"],
                     "segments_per_event": sum(segs) / n},
        "kernels": kernels,
        "accumulation": args.accum,
        "sensitivity_s": sens_s,
        "sensitivity_lors": n_lors,
        "plan_s": plan_s,
        "plan_bytes": plan_bytes,
        "plan_records": plan_records,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


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
            "seconds": dt}


if __name__ == "__main__":
    sys.exit(main())
