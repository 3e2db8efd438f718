"""Tile-planned backprojection (yrtmi_plan_create + *_planned, through the C
ABI) against the direct backprojection and the CPU oracle.

Bars (written in each test): planned vs oracle <= 1e-5 relative L2, as for
the direct path.  Planned vs direct: <= 2e-6 relative L2 on voxel sizes that
are short binary fractions (only the FP32 summation order differs: a walk
re-entered at a tile boundary computes the next plane as start + k * v,
which equals the direct walk's accumulated start + v + v + ... exactly
there); <= 1e-5 on other voxel sizes, where the two plane positions differ
by an ulp and segment ends at tile boundaries move by ~1e-5 mm.
"""
import numpy as np
import pytest
import torch

import oracle as O
import yrtmi
from helpers import oracle_image, oracle_proj, oracle_scanner, rel_l2
from yrtmi import synth

pytestmark = pytest.mark.gpu

TOL_ORACLE = 1e-5
TOL_DIRECT = 2e-6
TOL_DIRECT_NONDYADIC = 1e-5


def check(name, got, ref, tol):
    r = rel_l2(got, ref)
    print(f"[{name}] rel L2 = {r:.3e} (bar {tol:.0e})", flush=True)
    assert r <= tol, f"{name}: rel L2 {r:.3e} > {tol:.0e}"


def scanner():
    return yrtmi.Scanner("t", 96.0, 4.0, 4.0, 10.0, 180.0, 128, 24)


IMAGES = {
    # several tiles per axis, partial edge tiles
    "dyadic": lambda nt=1, off=(0.0, 0.0, 0.0): yrtmi.ImageParams(
        70, 66, 40, 4.0, 4.5, 2.25, off[0], off[1], off[2], nt),
    # voxel sizes that are not short binary fractions
    "nondyadic": lambda nt=1, off=(0.0, 0.0, 0.0): yrtmi.ImageParams(
        61, 57, 35, 5.1, 5.3, 2.7, off[0], off[1], off[2], nt),
}

CASES = [
    # (name, image, rays, tof, setup_fp64, lanes_per_event)
    ("r1", "dyadic", 1, None, 0, 1),
    ("r1_tof", "dyadic", 1, 250.0, 0, 1),
    ("r4_tof", "dyadic", 4, 250.0, 0, 4),
    ("r4_tof_fp64", "dyadic", 4, 250.0, 1, 4),
    ("r3_nondyadic", "nondyadic", 3, None, 0, 2),
    ("r4_tof_nondyadic_fp64", "nondyadic", 4, 300.0, 1, 4),
]


def setup(gpu, image, rays, tof, s64, lpe, nt=1, off=(0.0, 0.0, 0.0)):
    sc = scanner()
    im = IMAGES[image](nt, off)
    pp = yrtmi.ProjectorParams(sc, numRays=rays)
    pp.setupFP64 = bool(s64)
    if tof is not None:
        pp.addTOF(tof, 5)
    ctx = yrtmi.Context(sc)
    p = ctx.projector(pp)
    p.set_launch(lpe)
    return sc, im, pp, ctx, p


def events(sc, n, seed, tof, nt=1):
    rng = np.random.default_rng(seed)
    ev = synth.random_events(sc, n, rng, tof_fwhm_ps=tof)
    if nt > 1:
        ev["frame"] = rng.integers(-1, nt, n).astype(np.int32)
    return ev


def dev_lors(ev, gpu):
    g = lambda k, dt: None if k not in ev else torch.from_numpy(
        np.ascontiguousarray(ev[k]).astype(dt)).to(gpu)
    return yrtmi.LorSet.pairs(g("d1", np.int32), g("d2", np.int32), tof=g("tof", np.float32),
                              frame=g("frame", np.int32), motion_frame=g("motion", np.uint8))


def oracle_lors(sc, ev, xf=None):
    pos, ori = O.lut(oracle_scanner(sc))
    return O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev.get("tof"),
                             frame=ev.get("frame"), motion_frame=ev.get("motion"),
                             motion_xf=xf)


@pytest.mark.parametrize("name,image,rays,tof,s64,lpe", CASES, ids=[c[0] for c in CASES])
def test_back_planned(gpu, name, image, rays, tof, s64, lpe):
    sc, im, pp, ctx, p = setup(gpu, image, rays, tof, s64, lpe)
    n = 6000
    ev = events(sc, n, 101, tof)
    vals = np.random.default_rng(3).uniform(0.1, 1.0, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    vg = torch.from_numpy(vals).to(gpu)
    plan = ctx.plan(p, im, L)
    info = plan.info()
    print(f"{name}: plan {info}", flush=True)
    assert info["records"] >= n  # most LORs cross the image, each enters >= 1 tile
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, a, L, vg)
    ctx.back(p, im, b, L, vg, plan=plan)
    ctx.synchronize()
    ref = O.back(oracle_image(im), oracle_proj(pp), oracle_lors(sc, ev), vals,
                 dtype=np.float64 if s64 else np.float32)
    tol = TOL_DIRECT if image == "dyadic" else TOL_DIRECT_NONDYADIC
    check(f"{name}_planned_vs_direct", b.cpu().numpy(), a.cpu().numpy(), tol)
    check(f"{name}_planned_vs_oracle", b.cpu().numpy(), ref, TOL_ORACLE)
    # a plan is reusable: a second use adds the same image again
    ctx.back(p, im, b, L, vg, plan=plan)
    ctx.synchronize()
    check(f"{name}_reuse", b.cpu().numpy(), 2 * a.cpu().numpy(), tol)


def test_back_planned_frames_motion_offset(gpu):
    sc, im, pp, ctx, p = setup(gpu, "dyadic", 4, 250.0, 0, 4, nt=3, off=(3.0, -4.5, 2.0))
    rng = np.random.default_rng(12)
    xf = synth.rigid_transforms(rng, 4)
    ctx.upload_motion(xf)
    n = 6000
    ev = events(sc, n, 102, 250.0, nt=3)
    ev["motion"] = rng.integers(0, 4, n).astype(np.uint8)
    vals = rng.uniform(0.1, 1.0, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    vg = torch.from_numpy(vals).to(gpu)
    plan = ctx.plan(p, im, L)
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, a, L, vg)
    ctx.back(p, im, b, L, vg, plan=plan)
    ctx.synchronize()
    ref = O.back(oracle_image(im), oracle_proj(pp), oracle_lors(sc, ev, xf), vals)
    check("frames_planned_vs_direct", b.cpu().numpy(), a.cpu().numpy(), TOL_DIRECT)
    check("frames_planned_vs_oracle", b.cpu().numpy(), ref, TOL_ORACLE)


def test_plan_record_order_keeps_the_image(gpu):
    """A plan orders each chunk's records by length (kChunkSort) unless it
    is built for one use.  The records are the same, so the 64-bit
    fixed-point tile sums are too: with one chunk per tile (chunk_records
    16384, 32x32x8 tiles) the two plans' images are bit-identical.  The
    chunks hold several 2048-record sort windows, the last one partial."""
    sc, im, pp, ctx, p = setup(gpu, "dyadic", 4, 250.0, 0, 4)
    n = 20000
    ev = events(sc, n, 131, 250.0)
    vals = np.random.default_rng(13).uniform(0.1, 1.0, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    vg = torch.from_numpy(vals).to(gpu)
    ordered = ctx.plan(p, im, L, 16384, tile_shape=3)
    single = ctx.plan(p, im, L, 16384, tile_shape=3, single_use=True)
    info = ordered.info()
    assert info == single.info()
    assert info["records"] > 2048 * info["chunks"]  # several windows per chunk
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, a, L, vg, plan=ordered)
    ctx.back(p, im, b, L, vg, plan=single)
    ctx.synchronize()
    assert np.count_nonzero(a.cpu().numpy()) > 1000
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_back_planned_fixed_point_range(gpu):
    """Tile sums are accumulated in 64-bit fixed point scaled from the
    launch's values (terms out of range go to global atomics): values spread
    over 12 decades, or of both signs, match the direct FP32 backprojection
    (bars: 2e-6; 1e-5 for signed values, whose cancellation magnifies the
    direct path's own FP32 rounding) and the FP32 oracle (1e-5)."""
    sc, im, pp, ctx, p = setup(gpu, "dyadic", 4, 250.0, 0, 4)
    n = 6000
    ev = events(sc, n, 103, 250.0)
    rng = np.random.default_rng(5)
    L = dev_lors(ev, gpu)
    plan = ctx.plan(p, im, L)
    cases = (("decades", (10.0 ** rng.uniform(-6, 6, n)).astype(np.float32), TOL_DIRECT),
             ("signed", rng.uniform(-1.0, 1.0, n).astype(np.float32), 1e-5))
    for name, vals, tol in cases:
        vg = torch.from_numpy(vals).to(gpu)
        a = torch.zeros(im.shape, device=gpu)
        b = torch.zeros(im.shape, device=gpu)
        ctx.back(p, im, a, L, vg)
        ctx.back(p, im, b, L, vg, plan=plan)
        ctx.synchronize()
        check(f"fixed_{name}_planned_vs_direct", b.cpu().numpy(), a.cpu().numpy(), tol)
    ref = O.back(oracle_image(im), oracle_proj(pp), oracle_lors(sc, ev), vals)
    check("fixed_signed_planned_vs_oracle", b.cpu().numpy(), ref, 1e-5)


def test_back_planned_nonfinite_values(gpu):
    """A launch whose values are not all finite takes the FP32 tile path:
    the same voxels become non-finite as in the direct backprojection and
    the finite ones agree."""
    sc, im, pp, ctx, p = setup(gpu, "dyadic", 1, None, 0, 1)
    n = 4000
    ev = events(sc, n, 104, None)
    vals = np.random.default_rng(6).uniform(0.1, 1.0, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    plan = ctx.plan(p, im, L)
    # the LOR with the most records, so that the infinity reaches the image
    x1 = torch.ones(im.shape, device=gpu)
    fw = torch.zeros(n, device=gpu)
    ctx.forward(p, im, x1, L, fw)
    ctx.synchronize()
    vals[int(torch.argmax(fw).item())] = np.inf
    vg = torch.from_numpy(vals).to(gpu)
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, a, L, vg)
    ctx.back(p, im, b, L, vg, plan=plan)
    ctx.synchronize()
    a, b = a.cpu().numpy(), b.cpu().numpy()
    fa, fb = np.isfinite(a), np.isfinite(b)
    print(f"non-finite voxels: direct {int((~fa).sum())}, planned {int((~fb).sum())}")
    assert (~fa).sum() > 0
    assert np.array_equal(fa, fb)
    check("nonfinite_finite_part", b[fb], a[fa], TOL_DIRECT)


@pytest.mark.parametrize("rays,tof", [(1, None), (4, 250.0)])
def test_em_subset_planned(gpu, rays, tof):
    sc, im, pp, ctx, p = setup(gpu, "dyadic", rays, tof, 0, 4 if rays == 4 else 1)
    n = 8000
    ev = events(sc, n, 103, tof)
    rng = np.random.default_rng(4)
    x = rng.uniform(0.5, 2.0, im.shape).astype(np.float32)
    meas = rng.poisson(3.0, n).astype(np.float32)
    att = rng.uniform(0.2, 1.0, n).astype(np.float32)
    rnd = rng.uniform(0.0, 0.3, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    T = lambda a_: torch.from_numpy(a_).to(gpu)
    kw = dict(meas=T(meas), att=T(att), randoms=T(rnd), global_scale=0.9)
    plan = ctx.plan(p, im, L)
    u0 = torch.zeros(im.shape, device=gpu)
    u1 = torch.zeros(im.shape, device=gpu)
    ctx.em_subset(p, im, T(x), u0, L, **kw)
    ctx.em_subset(p, im, T(x), u1, L, plan=plan, **kw)
    ctx.synchronize()
    ref = np.zeros(im.shape, np.float32)
    O.em_subset(oracle_image(im), x, ref, oracle_proj(pp), oracle_lors(sc, ev), meas=meas,
                att=att, rnd=rnd, global_scale=0.9)
    check(f"em_r{rays}_planned_vs_direct", u1.cpu().numpy(), u0.cpu().numpy(), TOL_DIRECT)
    check(f"em_r{rays}_planned_vs_oracle", u1.cpu().numpy(), ref, TOL_ORACLE)


@pytest.mark.parametrize("rays,tof,lpe,s64", [(1, None, 1, 0), (4, 250.0, 4, 0),
                                              (4, 250.0, 2, 1)])
def test_em_ratio_planned_matches(gpu, rays, tof, lpe, s64):
    """em_ratio with the plan equals em_ratio without it, bit for bit."""
    sc, im, pp, ctx, p = setup(gpu, "dyadic", rays, tof, s64, lpe, nt=2)
    n = 6000
    ev = events(sc, n, 106, tof, nt=2)
    rng = np.random.default_rng(11)
    x = torch.from_numpy(rng.uniform(0.5, 2.0, im.shape).astype(np.float32)).to(gpu)
    meas = torch.from_numpy(rng.poisson(3.0, n).astype(np.float32)).to(gpu)
    L = dev_lors(ev, gpu)
    plan = ctx.plan(p, im, L)
    v0 = torch.zeros(n, device=gpu)
    v1 = torch.zeros(n, device=gpu)
    ctx.em_ratio(p, im, x, L, v0, meas=meas, global_scale=0.9)
    ctx.em_ratio(p, im, x, L, v1, meas=meas, global_scale=0.9, plan=plan)
    ctx.synchronize()
    # a third of the events are in frame -1 (skipped); many LORs miss the image
    assert torch.count_nonzero(v0).item() > n // 5
    assert torch.equal(v0, v1)


@pytest.mark.parametrize("att", [False, True])
def test_sensitivity_planned_histogram(gpu, att):
    sc = yrtmi.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, maxRingDiff=7)
    im = yrtmi.ImageParams(44, 40, 20, 3.0, 3.5, 2.0)
    pp = yrtmi.ProjectorParams(sc, numRays=2)
    ctx = yrtmi.Context(sc)
    p = ctx.projector(pp, with_tof=False)
    p.set_launch(2)
    osc = oracle_scanner(sc)
    S, s = 3, 2
    L = yrtmi.LorSet.histo_subset(ctx, S, s)
    mu = None
    if att:
        act = synth.cylinder_spheres(im, 40.0, 30.0, 1.0, [], supersample=1)
        mu = np.where(act > 0, np.float32(0.096), np.float32(0.0)).astype(np.float32)
    mug = torch.from_numpy(mu).to(gpu) if att else None
    # a single-use plan (records not ordered by length): the same image
    plan = ctx.plan(p, im, L, single_use=True)
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    kw = dict(att_im=im if att else None, att_img=mug)
    ctx.sensitivity(p, im, a, L, **kw)
    ctx.sensitivity(p, im, b, L, plan=plan, **kw)
    ctx.synchronize()
    pos, ori = O.lut(osc)
    osens = O.sensitivity_images(osc, oracle_image(im), oracle_proj(pp), S, False,
                                 att_im=oracle_image(im) if att else None, att_img=mu,
                                 hard_threshold=-1.0, pos_ori=(pos, ori))  # raw images
    check(f"sens_att{att}_planned_vs_direct", b.cpu().numpy(), a.cpu().numpy(), TOL_DIRECT)
    check(f"sens_att{att}_planned_vs_oracle", b.cpu().numpy(), osens[s], TOL_ORACLE)


def test_plan_misuse_rejected(gpu):
    sc, im, pp, ctx, p = setup(gpu, "dyadic", 1, None, 0, 1)
    ev = events(sc, 500, 104, None)
    L = dev_lors(ev, gpu)
    plan = ctx.plan(p, im, L)
    vals = torch.ones(500, device=gpu)
    img = torch.zeros(im.shape, device=gpu)
    other = yrtmi.ImageParams(70, 66, 40, 4.0, 4.5, 2.5)
    with pytest.raises(yrtmi.capi.YrtmiError, match="image geometry"):
        ctx.back(p, other, torch.zeros(other.shape, device=gpu), L, vals, plan=plan)
    with pytest.raises(yrtmi.capi.YrtmiError, match="LOR set"):
        ctx.back(p, im, img, L.slice(0, 400), vals, plan=plan)
    p2 = ctx.projector(pp)
    with pytest.raises(yrtmi.capi.YrtmiError, match="projector"):
        ctx.back(p2, im, img, L, vals, plan=plan)
    ptof = yrtmi.ProjectorParams(sc)
    ptof.addTOF(250.0, 5)
    pt = ctx.projector(ptof)
    Lt = dev_lors(events(sc, 500, 105, 250.0), gpu)
    plan_t = ctx.plan(pt, im, Lt)
    with pytest.raises(yrtmi.capi.YrtmiError, match="TOF off"):
        ctx.sensitivity(pt, im, img, Lt, plan=plan_t)
    pdd = ctx.projector(yrtmi.ProjectorParams(sc, projectorType="DD"))
    with pytest.raises(yrtmi.capi.YrtmiError, match="projector"):  # Siddon plan, DD projector
        ctx.back(pdd, im, img, L, vals, plan=plan)
    # more tiles (frames included) than any device's LDS holds bin counters for
    big = yrtmi.ImageParams(8, 8, 8, 1.0, 1.0, 1.0, nt=200000)
    with pytest.raises(yrtmi.capi.YrtmiError, match="image tiles"):
        ctx.plan(p, big, L)
