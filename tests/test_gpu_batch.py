"""Batching on the GPU: tile shapes, the persistent chunk-queue launch, the
autotuner, graph capture of an OSEM subset iteration, and many dynamic
frames of uneven event counts in one launch.

Bars: planned backprojections vs the direct one <= 2e-6 relative L2 (dyadic
voxel sizes: FP32 summation order only); graph replays vs the same calls run
eagerly <= 1e-6 relative L2 after every replay (atomic order only); the
batched frames vs frame-by-frame <= 1e-6.
"""
import numpy as np
import pytest
import torch

import yrtmi
from helpers import rel_l2
from yrtmi import synth

pytestmark = pytest.mark.gpu

TOL = 2e-6


def check(name, got, ref, tol):
    r = rel_l2(got, ref)
    print(f"[{name}] rel L2 = {r:.3e} (bar {tol:.0e})", flush=True)
    assert r <= tol, f"{name}: {r:.3e} > {tol:.0e}"


def setup(gpu, n=20000, nt=1, seed=2, stream=False):
    sc = yrtmi.Scanner("b", 96.0, 4.0, 4.0, 10.0, 180.0, 128, 24)
    im = yrtmi.ImageParams(70, 66, 40, 4.0, 4.5, 2.25, nt=nt)
    pp = yrtmi.ProjectorParams(sc, numRays=4)
    pp.addTOF(250.0, 5)
    ctx = yrtmi.Context(sc, stream=torch.cuda.Stream() if stream else None)
    proj = ctx.projector(pp)
    proj.set_launch(4)
    rng = np.random.default_rng(seed)
    ev = synth.random_events(sc, n, rng, tof_fwhm_ps=250.0)
    if nt > 1:
        ev["frame"] = rng.integers(0, nt, n).astype(np.int32)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(gpu)
    L = yrtmi.LorSet.pairs(T(ev["d1"], np.int32), T(ev["d2"], np.int32),
                           tof=T(ev["tof"], np.float32),
                           frame=T(ev["frame"], np.int32) if nt > 1 else None)
    return sc, im, pp, ctx, proj, ev, L, rng


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
@pytest.mark.parametrize("persistent", [0, 300])
def test_tile_shapes_and_persistent_queue(gpu, shape, persistent):
    sc, im, pp, ctx, proj, ev, L, rng = setup(gpu)
    vals = torch.from_numpy(rng.uniform(0.1, 1.0, L.n).astype(np.float32)).to(gpu)
    ref = torch.zeros(im.shape, device=gpu)
    ctx.back(proj, im, ref, L, vals)
    plan = ctx.plan(proj, im, L, tile_shape=shape, persistent=persistent)
    for k in range(2):  # the queue resets itself between launches
        out = torch.zeros(im.shape, device=gpu)
        ctx.back(proj, im, out, L, vals, plan=plan)
        ctx.synchronize()
        check(f"shape{shape}_pers{persistent}_run{k}", out.cpu().numpy(), ref.cpu().numpy(), TOL)


def test_autotune(gpu):
    sc, im, pp, ctx, proj, ev, L, rng = setup(gpu, n=40000)
    x = torch.from_numpy(rng.uniform(0.5, 1.5, im.shape).astype(np.float32)).to(gpu)
    t = ctx.autotune(proj, im, L, x)
    print("autotune:", t)
    assert t["lanes_per_event"] in (1, 2, 4, 8)
    assert t["tile_shape"] in (0, 1, 2, 3)
    assert 0 < t["ratio_ms"] == min(t["ratio_ms_lpe"])
    assert 0 < t["back_ms"] == min(ms for r in t["back_ms_shape"] for ms in r if ms > 0)
    # the tuned settings give the same EM update image
    u0 = torch.zeros(im.shape, device=gpu)
    ctx.em_subset(proj, im, x, u0, L)
    plan = ctx.plan(proj, im, L, tile_shape=t["tile_shape"], persistent=t["persistent"])
    u1 = torch.zeros(im.shape, device=gpu)
    ctx.em_subset(proj, im, x, u1, L, plan=plan)
    ctx.synchronize()
    check("tuned_em", u1.cpu().numpy(), u0.cpu().numpy(), TOL)


def _iteration(ctx, proj, im, x, u, sens, subsets, plans):
    for L, pl in zip(subsets, plans):
        ctx.fill(u, 0.0)
        ctx.em_subset(proj, im, x, u, L, plan=pl)
        ctx.em_update(x, u, sens, im.nvox)


def test_graph_replay_matches_eager(gpu):
    sc, im, pp, ctx, proj, ev, L, rng = setup(gpu, n=30000, stream=True)
    S = 3
    subsets = [L.slice(s * 10000, (s + 1) * 10000) for s in range(S)]
    plans = [ctx.plan(proj, im, Ls) for Ls in subsets]
    sens = torch.from_numpy(rng.uniform(0.5, 2.0, im.shape).astype(np.float32)).to(gpu)
    x0 = torch.from_numpy(rng.uniform(0.5, 1.5, im.shape).astype(np.float32)).to(gpu)
    with torch.cuda.stream(ctx.stream):
        xe = x0.clone()
        ue = torch.zeros_like(xe)
        xg = x0.clone()
        ug = torch.zeros_like(xg)
    ctx.synchronize()
    with ctx.graph() as g:
        _iteration(ctx, proj, im, xg, ug, sens, subsets, plans)
    for rep in range(3):
        g.launch()
        _iteration(ctx, proj, im, xe, ue, sens, subsets, plans)
        ctx.synchronize()
        # eager work of the program between replays
        scratch = (xe * 2.0).sum().item()
        assert np.isfinite(scratch)
        check(f"graph_replay_{rep}", xg.cpu().numpy(), xe.cpu().numpy(), 1e-6)


def test_graph_refuses_allocating_calls(gpu):
    sc, im, pp, ctx, proj, ev, L, rng = setup(gpu, n=2000, stream=True)
    with pytest.raises(yrtmi.capi.YrtmiError, match="recording"):
        with ctx.graph():
            ctx.plan(proj, im, L)
    # the context records and runs normally afterwards
    x = torch.ones(im.shape, device=gpu)
    out = torch.zeros(L.n, device=gpu)
    with ctx.graph() as g:
        ctx.forward(proj, im, x, L, out)
    g.launch()
    ctx.synchronize()
    ref = torch.zeros(L.n, device=gpu)
    ctx.forward(proj, im, x, L, ref)
    ctx.synchronize()
    check("graph_forward", out.cpu().numpy(), ref.cpu().numpy(), 1e-7)
    yrtmi.Context(sc)  # the null stream cannot record
    c0 = yrtmi.Context(sc)
    with pytest.raises(yrtmi.capi.YrtmiError, match="null stream"):
        with c0.graph():
            pass


def test_frames_in_one_launch(gpu):
    """Dynamic frames of uneven event counts as one multi-frame image and
    one LOR set (frame per event) equal the frame-by-frame computation."""
    F = 5
    counts = [500, 12000, 3000, 40000, 7000]
    sc = yrtmi.Scanner("b", 96.0, 4.0, 4.0, 10.0, 180.0, 128, 24)
    pp = yrtmi.ProjectorParams(sc, numRays=4)
    pp.addTOF(250.0, 5)
    ctx = yrtmi.Context(sc)
    proj = ctx.projector(pp)
    rng = np.random.default_rng(4)
    im1 = yrtmi.ImageParams(70, 66, 40, 4.0, 4.5, 2.25)
    imF = yrtmi.ImageParams(70, 66, 40, 4.0, 4.5, 2.25, nt=F)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(gpu)
    evs = [synth.random_events(sc, c, rng, tof_fwhm_ps=250.0) for c in counts]
    x = T(rng.uniform(0.5, 1.5, imF.shape), np.float32)
    # frame by frame
    u_ref = torch.zeros(imF.shape, device=gpu)
    for f, ev in enumerate(evs):
        Lf = yrtmi.LorSet.pairs(T(ev["d1"], np.int32), T(ev["d2"], np.int32),
                                tof=T(ev["tof"], np.float32))
        uf = torch.zeros(im1.shape, device=gpu)
        ctx.em_subset(proj, im1, x[f:f + 1].contiguous(), uf, Lf)
        u_ref[f] = uf[0]
    # all frames in one launch (tile-planned)
    cat = lambda k, dt: T(np.concatenate([ev[k] for ev in evs]), dt)
    frame = T(np.concatenate([np.full(c, f) for f, c in enumerate(counts)]), np.int32)
    LF = yrtmi.LorSet.pairs(cat("d1", np.int32), cat("d2", np.int32), tof=cat("tof", np.float32),
                            frame=frame)
    plan = ctx.plan(proj, imF, LF)
    u = torch.zeros(imF.shape, device=gpu)
    ctx.em_subset(proj, imF, x, u, LF, plan=plan)
    ctx.synchronize()
    check("frames_one_launch", u.cpu().numpy(), u_ref.cpu().numpy(), 1e-6)
