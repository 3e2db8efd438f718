"""Seeded GPU-resident workloads at the benchmark configurations.

Each builder returns a dict with the scanner, image geometry, context,
projector(s) and device tensors, with the events already permuted
subset-major (event i of the list belongs to subset i mod S, as the
reference's BinIteratorChronologicalInterleaved; after the permutation each
subset is one contiguous range `bounds[s]:bounds[s+1]`).
"""
from __future__ import annotations

import numpy as np
import torch

from . import core, synth


def listmode_tof_config(events=100_000_000, subsets=10, seed=2024, device="cuda:0",
                        rays=4, tof_fwhm_ps=250.0, tof_n_std=5, lpe=4, block=256,
                        grid=0, log=None):
    """64 rings x 512 crystals, R = 400 mm, 3.2 x 3.2 x 20 mm crystals;
    phantom: cylinder r = 150 mm, L = 180 mm at 1.0 plus 6 spheres r = 4..14
    mm at 5.0; image 200 x 200 x 128 at 2 mm."""
    dev = torch.device(device)
    sc = core.Scanner("mcd64", 64 * 3.2, 3.2, 3.2, 20.0, 400.0, 512, 64)
    im = core.ImageParams(200, 200, 128, 2.0, 2.0, 2.0)
    rng = np.random.default_rng(seed)
    spheres = synth.random_spheres_in_cylinder(rng, 150.0, 180.0,
                                               list(np.linspace(4.0, 14.0, 6)), 5.0)
    act = synth.cylinder_spheres(im, 150.0, 180.0, 1.0, spheres, supersample=2)
    ev = synth.sample_events(sc, im, act, events, seed=seed,
                             tof_fwhm_ps=tof_fwhm_ps, device=dev)
    n = events
    perm = torch.cat([torch.arange(s, n, subsets, device=dev) for s in range(subsets)])
    d1 = ev["d1"][perm].contiguous()
    d2 = ev["d2"][perm].contiguous()
    tof = ev["tof"][perm].contiguous()
    del ev, perm
    bounds = [0]
    for s in range(subsets):
        bounds.append(bounds[-1] + len(range(s, n, subsets)))
    ctx = core.Context(sc)
    pp = core.ProjectorParams(sc, numRays=rays)
    pp.addTOF(tof_fwhm_ps, tof_n_std)
    proj = ctx.projector(pp)
    proj.set_launch(lpe, block, grid)
    sub = [core.LorSet.pairs(d1[bounds[s]:bounds[s + 1]], d2[bounds[s]:bounds[s + 1]],
                             tof=tof[bounds[s]:bounds[s + 1]]) for s in range(subsets)]
    return dict(scanner=sc, image=im, ctx=ctx, params=pp, proj=proj, d1=d1, d2=d2,
                tof=tof, bounds=bounds, subsets=sub, activity=act, spheres=spheres,
                num_subsets=subsets, events=n)


def sensitivity_listmode(w, chunks=16, lpe=4, block=256, grid=0):
    """List-mode sensitivity image: backprojection of every scanner LOR (the
    uniform histogram, Corrector.cpp:152-157,228-241) without TOF, then the
    hard threshold (<= 1 -> 0; OSEM_CPU.cpp:168-172).  Returns (sens, n_lors)."""
    ctx, im = w["ctx"], w["image"]
    ps = ctx.projector(w["params"], with_tof=False)
    ps.set_launch(lpe, block, grid)
    sens = torch.zeros(im.shape, device=torch.device("cuda", ctx.device))
    n = 0
    for s in range(chunks):
        L = core.LorSet.histo_subset(ctx, chunks, s)
        ctx.sensitivity(ps, im, sens, L)
        n += L.n
    ctx.threshold(sens, sens, im.nvox, 1.0, 0.0, 0.0, 1.0, 0.0)
    return sens, n
