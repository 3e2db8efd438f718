"""Tile-planned distance-driven (DD) backprojection through the C ABI
(yrtmi_plan_create with a DD projector, then back / em_subset / sensitivity
with the plan) against the direct DD backprojection and the CPU oracle.

A DD plan's record is a (LOR, tile) pair; the tile kernel sweeps the LOR's
footprint over the tile's slices with the direct sweep's arithmetic, so the
voxel weights are bit-identical and only the summation order differs (the
tile sums in 64-bit fixed point, then one FP32 atomic per voxel).  Bars
(written in each test): planned vs direct <= 2e-6 relative L2; planned vs
oracle <= 1e-5 (projection-space) / 1e-4 (EM image), as for the direct path.
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
    "nondyadic": lambda nt=1, off=(0.0, 0.0, 0.0): yrtmi.ImageParams(
        61, 57, 35, 5.1, 5.3, 2.7, off[0], off[1], off[2], nt),
}


def setup(gpu, image, tof, s64=False, psf=False, nt=1, off=(0.0, 0.0, 0.0)):
    sc = scanner()
    im = IMAGES[image](nt, off)
    pp = yrtmi.ProjectorParams(sc, projectorType="DD")
    pp.setupFP64 = bool(s64)
    if tof is not None:
        pp.addTOF(tof, 5)
    if psf:
        k, s_step, ks = synth.gaussian_psf_row(2.0, k_spacing=0.25, num_s=3, s_step=40.0)
        k[1] *= 1.1  # rows differ so the row choice matters
        # an asymmetric row: the backprojection must use the flipped kernel
        k[2] = k[2][::-1] * np.linspace(0.8, 1.2, k.shape[1], dtype=np.float32)
        pp.psf_kernels, pp.psf_s_step, pp.psf_k_spacing = k, s_step, ks
    ctx = yrtmi.Context(sc)
    p = ctx.projector(pp)
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
                              frame=g("frame", np.int32))


def oracle_lors(sc, ev):
    pos, ori = O.lut(oracle_scanner(sc))
    return O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev.get("tof"),
                             frame=ev.get("frame"))


CASES = [
    # (name, image, tof, setup_fp64, psf, tile_shape)
    ("dd", "dyadic", None, False, False, 0),
    ("dd_tof", "dyadic", 250.0, False, False, 0),
    ("dd_tof_fp64", "dyadic", 250.0, True, False, 3),
    ("dd_psf", "dyadic", None, False, True, 3),
    ("dd_psf_tof", "nondyadic", 300.0, False, True, 0),
    ("dd_nondyadic_shape1", "nondyadic", None, False, False, 1),
    ("dd_tof_shape2_fp32_tile", "dyadic", 250.0, False, False, 2),
    # PSF tables staged in LDS beside the 128 KB tiles
    ("dd_psf_shape1", "nondyadic", None, False, True, 1),
    ("dd_psf_tof_shape2_fp32_tile", "dyadic", 250.0, False, True, 2),
]


def test_dd_psf_tables_beyond_lds(gpu):
    """A projection PSF of 48 rows of 57 taps (~34 KB per direction) is
    larger than the DD kernels stage in LDS (16 KB): the planned
    backprojection and EM ratio read it from global memory, and match the
    direct kernels and the oracle as the staged tables do."""
    sc = scanner()
    im = IMAGES["dyadic"]()
    pp = yrtmi.ProjectorParams(sc, projectorType="DD")
    k, s_step, ks = synth.gaussian_psf_row(2.0, k_spacing=0.125, num_s=48, s_step=5.0)
    assert k.shape == (48, 57)
    k = k * np.linspace(0.8, 1.2, 48, dtype=np.float32)[:, None]  # rows differ
    k[:, :20] *= 1.3  # asymmetric rows: the backprojection takes the mirrored ones
    pp.psf_kernels, pp.psf_s_step, pp.psf_k_spacing = k, s_step, ks
    ctx = yrtmi.Context(sc)
    p = ctx.projector(pp)
    n = 3000
    ev = events(sc, n, 151, None)
    vals = np.random.default_rng(15).uniform(0.1, 1.0, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    vg = torch.from_numpy(vals).to(gpu)
    plan = ctx.plan(p, im, L, tile_shape=0)
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, a, L, vg)
    ctx.back(p, im, b, L, vg, plan=plan)
    x = np.random.default_rng(16).uniform(0.5, 2.0, im.shape).astype(np.float32)
    xg = torch.from_numpy(x).to(gpu)
    mg = torch.from_numpy(np.random.default_rng(17).poisson(3.0, n).astype(np.float32)).to(gpu)
    v = torch.zeros(n, device=gpu)
    vp = torch.zeros(n, device=gpu)
    ctx.em_ratio(p, im, xg, L, v, meas=mg)
    ctx.em_ratio(p, im, xg, L, vp, meas=mg, plan=plan)
    ctx.synchronize()
    ref = O.back(oracle_image(im), oracle_proj(pp), oracle_lors(sc, ev), vals)
    assert np.count_nonzero(ref) > 1000
    check("dd_psf_global_planned_vs_direct", b.cpu().numpy(), a.cpu().numpy(), TOL_DIRECT)
    check("dd_psf_global_planned_vs_oracle", b.cpu().numpy(), ref, TOL_ORACLE)
    # the stored set-up gives the resolved LOR's set-up exactly
    np.testing.assert_array_equal(vp.cpu().numpy(), v.cpu().numpy())


def test_dd_plan_record_order_keeps_the_image(gpu):
    """DD records ordered by footprint voxels per slice (kChunkSort) or not
    (a single-use plan): one chunk per tile, so the fixed-point tile sums
    and the images are bit-identical."""
    sc, im, pp, ctx, p = setup(gpu, "dyadic", None, psf=True)
    n = 40000
    ev = events(sc, n, 141, None)
    vals = np.random.default_rng(14).uniform(0.1, 1.0, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    vg = torch.from_numpy(vals).to(gpu)
    ordered = ctx.plan(p, im, L, 1 << 20, tile_shape=0)
    single = ctx.plan(p, im, L, 1 << 20, tile_shape=0, single_use=True)
    info = ordered.info()
    assert info == single.info() and info["records"] > 2048 * info["chunks"]
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, a, L, vg, plan=ordered)
    ctx.back(p, im, b, L, vg, plan=single)
    ctx.synchronize()
    assert np.count_nonzero(a.cpu().numpy()) > 1000
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())


@pytest.mark.parametrize("name,image,tof,s64,psf,shape", CASES, ids=[c[0] for c in CASES])
def test_dd_back_planned(gpu, name, image, tof, s64, psf, shape):
    sc, im, pp, ctx, p = setup(gpu, image, tof, s64, psf)
    n = 4000
    ev = events(sc, n, 201, tof)
    vals = np.random.default_rng(5).uniform(0.1, 1.0, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    vg = torch.from_numpy(vals).to(gpu)
    plan = ctx.plan(p, im, L, tile_shape=shape)
    info = plan.info()
    print(f"{name}: plan {info}", flush=True)
    assert info["records"] >= n // 2
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, a, L, vg)
    ctx.back(p, im, b, L, vg, plan=plan)
    ctx.synchronize()
    # the FP32 oracle in both setup modes: DD keeps the reference FP32
    # footprint arithmetic, the FP64 setup only moves the slice range
    ref = O.back(oracle_image(im), oracle_proj(pp), oracle_lors(sc, ev), vals)
    assert np.count_nonzero(ref) > 1000
    check(f"{name}_planned_vs_direct", b.cpu().numpy(), a.cpu().numpy(), TOL_DIRECT)
    check(f"{name}_planned_vs_oracle", b.cpu().numpy(), ref, TOL_ORACLE)
    # reusable: a second use adds the same image again
    ctx.back(p, im, b, L, vg, plan=plan)
    ctx.synchronize()
    check(f"{name}_reuse", b.cpu().numpy(), 2 * a.cpu().numpy(), TOL_DIRECT)


def test_dd_back_planned_frames_offset(gpu):
    """Frames (-1 = skipped LOR) and an image offset: each frame's tiles take
    only that frame's LORs."""
    sc, im, pp, ctx, p = setup(gpu, "dyadic", 250.0, nt=3, off=(3.0, -4.5, 2.0))
    n = 5000
    ev = events(sc, n, 202, 250.0, nt=3)
    vals = np.random.default_rng(6).uniform(0.1, 1.0, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    vg = torch.from_numpy(vals).to(gpu)
    plan = ctx.plan(p, im, L, tile_shape=3)
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, a, L, vg)
    ctx.back(p, im, b, L, vg, plan=plan)
    ctx.synchronize()
    ref = O.back(oracle_image(im), oracle_proj(pp), oracle_lors(sc, ev), vals)
    check("dd_frames_planned_vs_direct", b.cpu().numpy(), a.cpu().numpy(), TOL_DIRECT)
    check("dd_frames_planned_vs_oracle", b.cpu().numpy(), ref, TOL_ORACLE)


def test_dd_back_planned_persistent_signed(gpu):
    """Persistent chunk queue, signed values spanning many binades (the
    fixed-point scale follows their median; the tail goes to global atomics)
    and zeros (skipped LORs)."""
    sc, im, pp, ctx, p = setup(gpu, "dyadic", None)
    n = 4000
    ev = events(sc, n, 203, None)
    rng = np.random.default_rng(7)
    vals = (rng.choice([-1.0, 1.0], n) * 10.0 ** rng.uniform(-3, 4, n)).astype(np.float32)
    vals[rng.random(n) < 0.2] = 0.0
    L = dev_lors(ev, gpu)
    vg = torch.from_numpy(vals).to(gpu)
    plan = ctx.plan(p, im, L, tile_shape=0, persistent=64)
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, a, L, vg)
    for _ in range(2):  # the queue resets itself between launches
        b.zero_()
        ctx.back(p, im, b, L, vg, plan=plan)
        ctx.synchronize()
        check("dd_persistent_signed_vs_direct", b.cpu().numpy(), a.cpu().numpy(), TOL_DIRECT)


@pytest.mark.parametrize("tof,psf", [(None, False), (250.0, True)])
def test_dd_em_subset_planned(gpu, tof, psf):
    sc, im, pp, ctx, p = setup(gpu, "dyadic", tof, psf=psf)
    n = 5000
    ev = events(sc, n, 204, tof)
    rng = np.random.default_rng(8)
    x = rng.uniform(0.5, 2.0, im.shape).astype(np.float32)
    meas = rng.poisson(3.0, n).astype(np.float32)
    att = rng.uniform(0.2, 1.0, n).astype(np.float32)
    rnd = rng.uniform(0.0, 0.3, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    T = lambda a_: torch.from_numpy(a_).to(gpu)
    kw = dict(meas=T(meas), att=T(att), randoms=T(rnd), global_scale=0.9)
    plan = ctx.plan(p, im, L, tile_shape=3)
    u0 = torch.zeros(im.shape, device=gpu)
    u1 = torch.zeros(im.shape, device=gpu)
    ctx.em_subset(p, im, T(x), u0, L, **kw)
    ctx.em_subset(p, im, T(x), u1, L, plan=plan, **kw)
    ctx.synchronize()
    ref = np.zeros(im.shape, np.float32)
    O.em_subset(oracle_image(im), x, ref, oracle_proj(pp), oracle_lors(sc, ev), meas=meas,
                att=att, rnd=rnd, global_scale=0.9)
    check(f"dd_em_tof{tof}_psf{psf}_planned_vs_direct", u1.cpu().numpy(), u0.cpu().numpy(),
          TOL_DIRECT)
    check(f"dd_em_tof{tof}_psf{psf}_planned_vs_oracle", u1.cpu().numpy(), ref, 1e-4)


def test_dd_em_ratio_matches_forward(gpu):
    """em_ratio with a DD projector: m / (forward projection) per LOR."""
    sc, im, pp, ctx, p = setup(gpu, "dyadic", 250.0)
    n = 3000
    ev = events(sc, n, 205, 250.0)
    rng = np.random.default_rng(9)
    x = rng.uniform(0.5, 2.0, im.shape).astype(np.float32)
    meas = rng.poisson(3.0, n).astype(np.float32)
    L = dev_lors(ev, gpu)
    xg = torch.from_numpy(x).to(gpu)
    fwd = torch.zeros(n, device=gpu)
    ctx.forward(p, im, xg, L, fwd)
    v = torch.zeros(n, device=gpu)
    mg = torch.from_numpy(meas).to(gpu)
    ctx.em_ratio(p, im, xg, L, v, meas=mg)
    # from the plan's table of stored set-ups: the same values
    vp = torch.zeros(n, device=gpu)
    ctx.em_ratio(p, im, xg, L, vp, meas=mg, plan=ctx.plan(p, im, L))
    ctx.synchronize()
    f = fwd.cpu().numpy()
    want = np.where((meas != 0) & (np.abs(f) > 1e-35), meas / np.where(f == 0, 1, f), 0.0)
    np.testing.assert_array_equal(v.cpu().numpy(), want.astype(np.float32))
    np.testing.assert_array_equal(vp.cpu().numpy(), want.astype(np.float32))


def test_dd_sensitivity_planned_histogram(gpu):
    sc = yrtmi.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, maxRingDiff=7)
    im = yrtmi.ImageParams(44, 40, 20, 3.0, 3.5, 2.0)
    pp = yrtmi.ProjectorParams(sc, projectorType="DD")
    ctx = yrtmi.Context(sc)
    p = ctx.projector(pp, with_tof=False)
    osc = oracle_scanner(sc)
    S, s = 3, 1
    L = yrtmi.LorSet.histo_subset(ctx, S, s)
    act = synth.cylinder_spheres(im, 40.0, 30.0, 1.0, [], supersample=1)
    mu = np.where(act > 0, np.float32(0.096), np.float32(0.0)).astype(np.float32)
    mug = torch.from_numpy(mu).to(gpu)
    plan = ctx.plan(p, im, L, tile_shape=0)
    a = torch.zeros(im.shape, device=gpu)
    b = torch.zeros(im.shape, device=gpu)
    ctx.sensitivity(p, im, a, L, att_im=im, att_img=mug)
    ctx.sensitivity(p, im, b, L, att_im=im, att_img=mug, plan=plan)
    ctx.synchronize()
    pos, ori = O.lut(osc)
    osens = O.sensitivity_images(osc, oracle_image(im), oracle_proj(pp), S, False,
                                 att_im=oracle_image(im), att_img=mu, hard_threshold=-1.0,
                                 pos_ori=(pos, ori))  # raw images
    check("dd_sens_planned_vs_direct", b.cpu().numpy(), a.cpu().numpy(), TOL_DIRECT)
    check("dd_sens_planned_vs_oracle", b.cpu().numpy(), osens[s], TOL_ORACLE)
