"""Locality order of LOR sets (yrtmi_lor_order / yrtmi_permute) on the GPU.

The order has no counterpart in the reference; what it must preserve is
that results do not depend on it:
- the output is a permutation of 0..n-1, with skipped LORs (frame out of
  range) last;
- forward projection commutes with the permutation exactly (each LOR's
  value is computed on its own);
- the EM update image of a reordered subset equals the original's to FP32
  summation order (<= 1e-6 relative L2, direct and tile-planned);
- the host OSEM with ordered list-mode subsets equals it without (<= 1e-5
  relative L2 after 2 iterations x 3 subsets).
"""
import numpy as np
import pytest
import torch

import pyyrtmi as Y
import yrtmi
from helpers import rel_l2
from yrtmi import synth

pytestmark = pytest.mark.gpu


def check(name, got, ref, tol):
    r = rel_l2(got, ref)
    print(f"[{name}] rel L2 = {r:.3e} (bar {tol:.0e})", flush=True)
    assert r <= tol, f"{name}: {r:.3e} > {tol:.0e}"


def setup(gpu, n=30000, nt=1, seed=5):
    sc = yrtmi.Scanner("o", 96.0, 4.0, 4.0, 10.0, 180.0, 128, 24)
    im = yrtmi.ImageParams(64, 64, 40, 4.0, 4.0, 2.5, nt=nt)
    pp = yrtmi.ProjectorParams(sc, numRays=4)
    pp.addTOF(250.0, 5)
    ctx = yrtmi.Context(sc)
    proj = ctx.projector(pp)
    rng = np.random.default_rng(seed)
    ev = synth.random_events(sc, n, rng, tof_fwhm_ps=250.0)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(gpu)
    frame = None
    if nt > 1:
        frame = T(rng.integers(0, nt + 1, n), np.int32)  # frame nt: skipped
    L = yrtmi.LorSet.pairs(T(ev["d1"], np.int32), T(ev["d2"], np.int32),
                           tof=T(ev["tof"], np.float32), frame=frame)
    return sc, im, pp, ctx, proj, L, rng


def test_order_is_a_permutation_with_skipped_last(gpu):
    sc, im, pp, ctx, proj, L, rng = setup(gpu, nt=3)
    perm = ctx.lor_order(im, L)
    ctx.synchronize()
    p = perm.cpu().numpy().astype(np.int64)
    assert np.array_equal(np.sort(p), np.arange(L.n))
    skipped = (L.frame.cpu().numpy() >= im.nt)[p]
    k = int(skipped.sum())
    assert k > 0 and skipped[L.n - k:].all() and not skipped[:L.n - k].any()
    # frames are the most significant part of the key
    f = L.frame.cpu().numpy()[p[:L.n - k]]
    assert (np.diff(f) >= 0).all()
    # deterministic
    perm2 = ctx.lor_order(im, L)
    ctx.synchronize()
    assert torch.equal(perm, perm2)


def test_forward_commutes_with_permutation(gpu):
    sc, im, pp, ctx, proj, L, rng = setup(gpu)
    x = torch.from_numpy(rng.uniform(0.5, 1.5, im.shape).astype(np.float32)).to(gpu)
    y = torch.zeros(L.n, device=gpu)
    ctx.forward(proj, im, x, L, y)
    perm = ctx.lor_order(im, L)
    Lp = L.permuted(ctx, perm)
    yp = torch.zeros(L.n, device=gpu)
    ctx.forward(proj, im, x, Lp, yp)
    ctx.synchronize()
    assert torch.equal(yp, y[perm.long()])


def test_em_update_independent_of_order(gpu):
    sc, im, pp, ctx, proj, L, rng = setup(gpu)
    x = torch.from_numpy(rng.uniform(0.5, 1.5, im.shape).astype(np.float32)).to(gpu)
    Lp = L.permuted(ctx, ctx.lor_order(im, L))
    u0 = torch.zeros(im.shape, device=gpu)
    ctx.em_subset(proj, im, x, u0, L)
    u1 = torch.zeros(im.shape, device=gpu)
    ctx.em_subset(proj, im, x, u1, Lp)
    u2 = torch.zeros(im.shape, device=gpu)
    ctx.em_subset(proj, im, x, u2, Lp, plan=ctx.plan(proj, im, Lp))
    ctx.synchronize()
    check("em_ordered_direct", u1.cpu().numpy(), u0.cpu().numpy(), 1e-6)
    check("em_ordered_planned", u2.cpu().numpy(), u0.cpu().numpy(), 2e-6)


def test_permute_element_sizes(gpu):
    sc, im, pp, ctx, proj, L, rng = setup(gpu, n=5000)
    perm = ctx.lor_order(im, L)
    pl = perm.long()
    for dt in (torch.uint8, torch.int16, torch.int32, torch.float64):
        src = torch.from_numpy(rng.integers(0, 100, L.n)).to(gpu).to(dt)
        out = ctx.permute(src, perm)
        ctx.synchronize()
        assert torch.equal(out, src[pl]), dt
    with pytest.raises(yrtmi.capi.YrtmiError, match="elem_bytes"):
        yrtmi.capi.check(yrtmi.capi.lib().yrtmi_permute(ctx.h, src.data_ptr(), 3,
                                                         perm.data_ptr(), 10, out.data_ptr()))


def test_host_osem_order_on_off(gpu):
    a = ("lm", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    sc, ysc = yrtmi.Scanner(*a), Y.Scanner(*a, 1, 15, 1, 1)
    im = yrtmi.ImageParams(40, 40, 24, 4.0, 4.0, 2.75)
    yim = Y.ImageParams.fromParams(40, 40, 24, 4.0, 4.0, 2.75, 1)
    act = synth.cylinder_spheres(im, 60.0, 50.0, 1.0, [(10.0, -5.0, 3.0, 10.0, 4.0)])
    ev = synth.sample_events(sc, im, act, 60000, seed=3, tof_fwhm_ps=250.0)
    out = []
    for order in (False, True):
        osem = Y.createOSEM(ysc)
        c = osem.getContext()
        lm = Y.ListModeDeviceOwned(c, ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32),
                                   tof=ev["tof"].astype(np.float32))
        osem.setImageParams(yim)
        osem.setNumRays(4)
        osem.addTOF(250.0, 5)
        osem.useLorOrder = order
        osem.num_OSEM_subsets = 3
        osem.num_MLEM_iterations = 2
        osem.setDataInput(lm)
        osem.generateSensitivityImages()
        out.append(osem.reconstruct().transferToHostMemory())
    check("host_osem_order", out[1], out[0], 1e-5)
