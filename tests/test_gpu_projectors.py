"""GPU parity of the projectors (through the C ABI) against the CPU oracle.

Bars (written in each test): projections <= 1e-5 relative L2 vs the FP32
oracle; backprojected images <= 1e-5 relative L2 (atomic accumulation order
differs from the oracle's sequential order).  The FP64 oracle error of both
is printed for the record.
"""
import numpy as np
import pytest
import torch

import oracle as O
import yrtmi
from helpers import oracle_image, oracle_proj, oracle_scanner, rel_l2
from yrtmi import synth

pytestmark = pytest.mark.gpu

TOL_PROJ = 1e-5
TOL_IMG = 1e-5


def check(name, got, ref, tol):
    r = rel_l2(got, ref)
    print(f"[{name}] rel L2 = {r:.3e} (bar {tol:.0e})", flush=True)
    assert r <= tol, f"{name}: rel L2 {r:.3e} > {tol:.0e}"


def small_scanner():
    # 16 rings x 96 crystals, radius 150 mm, 4x4x10 mm crystals
    return yrtmi.Scanner("small", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)


def small_image(nt=1, off=(0.0, 0.0, 0.0)):
    return yrtmi.ImageParams(40, 36, 16, 5.0, 5.5, 4.0, off[0], off[1], off[2], nt)


def make_events(sc, n, seed, tof=None, nt=1):
    rng = np.random.default_rng(seed)
    ev = synth.random_events(sc, n, rng, tof_fwhm_ps=tof)
    if nt > 1:
        ev["frame"] = rng.integers(-1, nt, n).astype(np.int32)
    return ev


def to_dev(ev, dev):
    g = lambda k, dt: None if k not in ev else torch.from_numpy(np.ascontiguousarray(ev[k]).astype(dt)).to(dev)
    return yrtmi.LorSet.pairs(g("d1", np.int32), g("d2", np.int32), tof=g("tof", np.float32),
                              frame=g("frame", np.int32))


def oracle_lors(sc, ev, motion_xf=None):
    pos, ori = O.lut(oracle_scanner(sc))
    return O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev.get("tof"),
                             frame=ev.get("frame"),
                             motion_frame=ev.get("motion"), motion_xf=motion_xf)


CASES = [
    # (name, projector, rays, tof_fwhm, lanes_per_event)
    ("siddon", "Siddon", 1, None, 1),
    ("siddon_tof", "Siddon", 1, 250.0, 1),
    ("mr4_lpe1", "Siddon", 4, None, 1),
    ("mr4_lpe4", "Siddon", 4, None, 4),
    ("mr4_tof_lpe4", "Siddon", 4, 250.0, 4),
    ("mr6_tof_lpe2", "Siddon", 6, 300.0, 2),
    ("mr3_lpe8", "Siddon", 3, None, 8),
    ("dd", "DD", 1, None, 1),
    ("dd_tof", "DD", 1, 250.0, 1),
]


def _setup(gpu, proj, rays, tof, lpe, nt=1, off=(0.0, 0.0, 0.0), psf=False):
    sc = small_scanner()
    im = small_image(nt, off)
    pp = yrtmi.ProjectorParams(sc, projectorType=proj, numRays=rays)
    if tof is not None:
        pp.addTOF(tof, 5)
    if psf:
        k, s_step, ks = synth.gaussian_psf_row(2.0, k_spacing=0.25, num_s=3, s_step=40.0)
        k[1] *= 1.1  # rows differ so the row choice matters
        k[2] = k[2][::-1] * np.linspace(0.8, 1.2, k.shape[1], dtype=np.float32)
        pp.psf_kernels, pp.psf_s_step, pp.psf_k_spacing = k, s_step, ks
    ctx = yrtmi.Context(sc)
    p = ctx.projector(pp)
    p.set_launch(lpe)
    return sc, im, pp, ctx, p


@pytest.mark.parametrize("name,proj,rays,tof,lpe", CASES, ids=[c[0] for c in CASES])
def test_forward_back(gpu, name, proj, rays, tof, lpe):
    sc, im, pp, ctx, p = _setup(gpu, proj, rays, tof, lpe)
    ev = make_events(sc, 3000, seed=11, tof=tof)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.0, 2.0, im.shape).astype(np.float32)
    vals = rng.uniform(0.1, 1.0, 3000).astype(np.float32)
    L = to_dev(ev, gpu)
    out = torch.zeros(3000, device=gpu)
    ctx.forward(p, im, torch.from_numpy(x).to(gpu), L, out)
    img = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, img, L, torch.from_numpy(vals).to(gpu))
    ctx.synchronize()
    OL = oracle_lors(sc, ev)
    po = oracle_proj(pp)
    oi = oracle_image(im)
    f32 = O.forward(oi, x, po, OL)
    f64 = O.forward(oi, x.astype(np.float64), po, OL, dtype=np.float64)
    b32 = O.back(oi, po, OL, vals)
    b64 = O.back(oi, po, OL, vals.astype(np.float64), dtype=np.float64)
    g_f = out.cpu().numpy()
    g_b = img.cpu().numpy()
    print(f"{name}: fwd gpu-vs-f32 {rel_l2(g_f, f32):.2e} gpu-vs-f64 {rel_l2(g_f, f64):.2e} "
          f"f32-vs-f64 {rel_l2(f32, f64):.2e}; back gpu-vs-f32 {rel_l2(g_b, b32):.2e} "
          f"gpu-vs-f64 {rel_l2(g_b, b64):.2e} f32-vs-f64 {rel_l2(b32, b64):.2e}")
    assert np.count_nonzero(f32) > 100
    check("g_f", g_f, f32, TOL_PROJ)
    check("g_b", g_b, b32, TOL_IMG)


@pytest.mark.parametrize("tof", [None, 250.0])
def test_dd_psf(gpu, tof):
    sc, im, pp, ctx, p = _setup(gpu, "DD", 1, tof, 1, psf=True)
    ev = make_events(sc, 2000, seed=21, tof=tof)
    rng = np.random.default_rng(6)
    x = rng.uniform(0.0, 2.0, im.shape).astype(np.float32)
    vals = rng.uniform(0.1, 1.0, 2000).astype(np.float32)
    L = to_dev(ev, gpu)
    out = torch.zeros(2000, device=gpu)
    ctx.forward(p, im, torch.from_numpy(x).to(gpu), L, out)
    img = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, img, L, torch.from_numpy(vals).to(gpu))
    ctx.synchronize()
    OL = oracle_lors(sc, ev)
    po = oracle_proj(pp)
    oi = oracle_image(im)
    f32 = O.forward(oi, x, po, OL)
    b32 = O.back(oi, po, OL, vals)
    check(f"dd_psf_fwd_{tof}", out.cpu().numpy(), f32, TOL_PROJ)
    check(f"dd_psf_back_{tof}", img.cpu().numpy(), b32, TOL_IMG)


def test_offset_dynamic_frames(gpu):
    sc, im, pp, ctx, p = _setup(gpu, "Siddon", 4, 250.0, 4, nt=3, off=(3.0, -4.5, 2.0))
    ev = make_events(sc, 3000, seed=31, tof=250.0, nt=3)
    rng = np.random.default_rng(7)
    x = rng.uniform(0.0, 2.0, im.shape).astype(np.float32)
    vals = rng.uniform(0.1, 1.0, 3000).astype(np.float32)
    L = to_dev(ev, gpu)
    out = torch.zeros(3000, device=gpu)
    ctx.forward(p, im, torch.from_numpy(x).to(gpu), L, out)
    img = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, img, L, torch.from_numpy(vals).to(gpu))
    ctx.synchronize()
    OL = oracle_lors(sc, ev)
    po = oracle_proj(pp)
    oi = oracle_image(im)
    f32 = O.forward(oi, x, po, OL)
    b32 = O.back(oi, po, OL, vals)
    assert (f32[ev["frame"] < 0] == 0).all()
    check("out.cpu().numpy()", out.cpu().numpy(), f32, TOL_PROJ)
    check("img.cpu().numpy()", img.cpu().numpy(), b32, TOL_IMG)


def test_motion(gpu):
    sc, im, pp, ctx, p = _setup(gpu, "Siddon", 1, 250.0, 1)
    rng = np.random.default_rng(8)
    xf = synth.rigid_transforms(rng, 4)
    ctx.upload_motion(xf)
    ev = make_events(sc, 2000, seed=41, tof=250.0)
    ev["motion"] = rng.integers(0, 4, 2000).astype(np.uint8)
    x = rng.uniform(0.0, 2.0, im.shape).astype(np.float32)
    L = to_dev(ev, gpu)
    L.motion_frame = torch.from_numpy(ev["motion"]).to(gpu)
    out = torch.zeros(2000, device=gpu)
    ctx.forward(p, im, torch.from_numpy(x).to(gpu), L, out)
    ctx.synchronize()
    OL = oracle_lors(sc, ev, motion_xf=xf)
    f32 = O.forward(oracle_image(im), x, oracle_proj(pp), OL)
    check("out.cpu().numpy()", out.cpu().numpy(), f32, TOL_PROJ)


def test_histogram_sources(gpu):
    """Bins decoded on the device (bin list and subset iterator) give the
    same LORs as the oracle's Histogram3D mapping."""
    sc = yrtmi.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, maxRingDiff=7)
    im = yrtmi.ImageParams(32, 32, 12, 4.0, 4.0, 4.0)
    pp = yrtmi.ProjectorParams(sc)
    ctx = yrtmi.Context(sc)
    p = ctx.projector(pp)
    osc = oracle_scanner(sc)
    assert ctx.histo_dims == O.histo_dims(osc)
    rng = np.random.default_rng(9)
    x = rng.uniform(0.0, 2.0, im.shape).astype(np.float32)
    xg = torch.from_numpy(x).to(gpu)
    S, s = 4, 1
    L = yrtmi.LorSet.histo_subset(ctx, S, s)
    bins = O.histo_subset_bins(osc, S, s)
    assert L.n == bins.size
    out = torch.zeros(L.n, device=gpu)
    ctx.forward(p, im, xg, L, out)
    d1, d2 = O.histo_pairs(osc, bins)
    pos, ori = O.lut(osc)
    f32 = O.forward(oracle_image(im), x, oracle_proj(pp), O.lors_from_pairs(pos, ori, d1, d2))
    # explicit bin list (shuffled subset)
    perm = rng.permutation(bins.size)[:5000]
    Lb = yrtmi.LorSet.bins_(torch.from_numpy(bins[perm].astype(np.int32)).to(gpu))
    outb = torch.zeros(Lb.n, device=gpu)
    ctx.forward(p, im, xg, Lb, outb)
    ctx.synchronize()
    check("out.cpu().numpy()", out.cpu().numpy(), f32, TOL_PROJ)
    check("outb.cpu().numpy()", outb.cpu().numpy(), f32[perm], TOL_PROJ)


@pytest.mark.parametrize("proj,rays,tof", [("Siddon", 4, 250.0), ("Siddon", 1, None), ("DD", 1, 250.0)])
def test_em_subset(gpu, proj, rays, tof):
    sc, im, pp, ctx, p = _setup(gpu, proj, rays, tof, 4 if rays == 4 else 1)
    ev = make_events(sc, 4000, seed=51, tof=tof)
    rng = np.random.default_rng(10)
    x = rng.uniform(0.5, 2.0, im.shape).astype(np.float32)
    meas = rng.poisson(2.0, 4000).astype(np.float32)
    att = rng.uniform(0.2, 1.0, 4000).astype(np.float32)
    rnd = rng.uniform(0.0, 0.1, 4000).astype(np.float32)
    L = to_dev(ev, gpu)
    u = torch.zeros(im.shape, device=gpu)
    T = lambda a: torch.from_numpy(a).to(gpu)
    ctx.em_subset(p, im, T(x), u, L, meas=T(meas), att=T(att), randoms=T(rnd), global_scale=1.3)
    ctx.synchronize()
    uo = np.zeros(im.shape, np.float32)
    O.em_subset(oracle_image(im), x, uo, oracle_proj(pp), oracle_lors(sc, ev), meas=meas,
                att=att, rnd=rnd, global_scale=1.3)
    check("u.cpu().numpy()", u.cpu().numpy(), uo, TOL_IMG)


def test_count_segments(gpu):
    sc, im, pp, ctx, p = _setup(gpu, "Siddon", 4, 250.0, 4)
    ev = make_events(sc, 3000, seed=61, tof=250.0)
    n = ctx.count_segments(p, im, to_dev(ev, gpu))
    c = O.count_segments(oracle_image(im), oracle_proj(pp), oracle_lors(sc, ev))
    # identical traversal except for FP64-vs-FP32 setup on grazing rays
    assert abs(n - int(c.sum())) <= max(3, 1e-4 * c.sum())


def test_adjoint_gpu(gpu):
    """<A x, y> = <x, A^T y> with the GPU forward and back kernels."""
    sc, im, pp, ctx, p = _setup(gpu, "Siddon", 4, 250.0, 4)
    ev = make_events(sc, 5000, seed=71, tof=250.0)
    rng = np.random.default_rng(12)
    x = torch.from_numpy(rng.uniform(0, 1, im.shape).astype(np.float32)).to(gpu)
    y = torch.from_numpy(rng.uniform(0, 1, 5000).astype(np.float32)).to(gpu)
    L = to_dev(ev, gpu)
    ax = torch.zeros(5000, device=gpu)
    ctx.forward(p, im, x, L, ax)
    aty = torch.zeros(im.shape, device=gpu)
    ctx.back(p, im, aty, L, y)
    ctx.synchronize()
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_sensitivity_and_update(gpu):
    sc = yrtmi.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, maxRingDiff=7)
    im = yrtmi.ImageParams(32, 32, 12, 4.0, 4.0, 4.0)
    pp = yrtmi.ProjectorParams(sc)
    ctx = yrtmi.Context(sc)
    p = ctx.projector(pp, with_tof=False)
    osc = oracle_scanner(sc)
    rng = np.random.default_rng(13)
    mu = (rng.uniform(0, 1, im.shape[1:]) * 0.096).astype(np.float32)
    img = torch.zeros(im.shape, device=gpu)
    L = yrtmi.LorSet.histo_subset(ctx, 2, 1)
    ctx.sensitivity(p, im, img, L, att_im=im, att_img=torch.from_numpy(mu).to(gpu))
    ctx.synchronize()
    bins = O.histo_subset_bins(osc, 2, 1)
    d1, d2 = O.histo_pairs(osc, bins)
    pos, ori = O.lut(osc)
    OL = O.lors_from_pairs(pos, ori, d1, d2)
    acf = O.attenuation_factors(oracle_image(im), mu, OL)
    so = O.back(oracle_image(im), oracle_proj(pp, with_tof=False), OL, acf)
    check("img.cpu().numpy()", img.cpu().numpy(), so, TOL_IMG)
    # per-LOR attenuation factors
    Lb = yrtmi.LorSet.bins_(torch.from_numpy(bins.astype(np.int32)).to(gpu))
    a = torch.zeros(Lb.n, device=gpu)
    ctx.attenuation_factors(im, torch.from_numpy(mu).to(gpu), Lb, a)
    ctx.synchronize()
    check("a.cpu().numpy()", a.cpu().numpy(), acf, TOL_PROJ)
    # EM update / threshold, elementwise: exact
    x = rng.uniform(0, 2, im.shape).astype(np.float32)
    u = rng.uniform(0, 2, im.shape).astype(np.float32)
    s = rng.uniform(-0.5, 2, im.shape[1:]).astype(np.float32)
    xg = torch.from_numpy(x).to(gpu)
    ctx.em_update(xg, torch.from_numpy(u).to(gpu), torch.from_numpy(s).to(gpu), im.nvox)
    xo = O.em_update(x.copy(), u, s)
    ctx.threshold(xg, torch.from_numpy(s).to(gpu), im.nvox, 0.3, 0.0, 0.0, 2.0, 0.5)
    xo = O.apply_threshold(xo, s, 0.3, 0.0, 0.0, 2.0, 0.5)
    ctx.synchronize()
    np.testing.assert_allclose(xg.cpu().numpy(), xo.reshape(im.shape), rtol=2e-7, atol=0)


@pytest.mark.parametrize("proj,rays,tof", [("Siddon", 1, 250.0), ("Siddon", 4, 250.0),
                                           ("Siddon", 4, None), ("DD", 1, 250.0)])
def test_setup_fp64_vs_fp64_oracle(gpu, proj, rays, tof):
    """Accuracy mode (FP64 ray-plane setup) against the FP64 oracle.  The
    reference FP32 order is off by up to several % on rays grazing a voxel
    plane (e.g. x ~ 1e-6 mm along x = 0); FP64 setup is not."""
    sc, im, pp, ctx, p = _setup(gpu, proj, rays, tof, 4 if rays == 4 else 1)
    pp.setupFP64 = True
    p = ctx.projector(pp)
    p.set_launch(4 if rays == 4 else 1)
    ev = make_events(sc, 3000, seed=11, tof=tof)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.0, 2.0, im.shape).astype(np.float32)
    L = to_dev(ev, gpu)
    out = torch.zeros(3000, device=gpu)
    ctx.forward(p, im, torch.from_numpy(x).to(gpu), L, out)
    ctx.synchronize()
    OL = oracle_lors(sc, ev)
    po = oracle_proj(pp)
    f64 = O.forward(oracle_image(im), x.astype(np.float64), po, OL, dtype=np.float64)
    f32 = O.forward(oracle_image(im), x, po, OL)
    print(f"fp64-setup gpu vs f64 {rel_l2(out.cpu().numpy(), f64):.2e}; "
          f"reference f32 vs f64 {rel_l2(f32, f64):.2e}")
    if proj == "DD":
        # DD's error is the reference FP32 footprint arithmetic (kept in both
        # modes); the FP64 setup only moves the slice range
        check(f"fp64setup_{proj}{rays}_{tof}", out.cpu().numpy(), f32, TOL_PROJ)
    else:
        check(f"fp64setup_{proj}{rays}_{tof}", out.cpu().numpy(), f64, TOL_PROJ)
