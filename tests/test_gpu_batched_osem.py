"""The C++ host OSEM (pyyrtmi.OSEM) in its batched and budgeted forms, on
the GPU, against the oracle's restatement of the reference CPU OSEM flow:

* histogram frames reconstructed together (one launch per subset for all
  frames) with the distance-driven projector and the projection-space PSF:
  each frame equals the oracle's reconstruction of that frame's histogram;
* the 24-ring x 256-crystal histogram configuration at full size (max ring
  difference 23, 96x96x48 @ 2 mm, attenuation, 5 iterations x 8 subsets);
* a device-memory budget for the tile plans: subsets beyond it run the
  unplanned kernels and the images do not change;
* graph-replayed iterations and autotuned launch shapes against the
  call-by-call defaults.

Bars: sensitivity images <= 1e-5 relative L2; images after the iteration
count <= 1e-4 (atomic-order nondeterminism included)."""
import numpy as np
import pytest

import oracle as O
import pyyrtmi as Y
import yrtmi
from helpers import oracle_image, oracle_scanner, rel_l2
from yrtmi import synth

pytestmark = pytest.mark.gpu

TOL_SENS = 1e-5
TOL_IMG = 1e-4


def check(name, got, ref, tol):
    r = rel_l2(got, ref)
    print(f"[{name}] rel L2 = {r:.3e} (bar {tol:.0e})", flush=True)
    assert r <= tol, f"{name}: {r:.3e} > {tol:.0e}"


def histo_setup(a, mrd, imdims):
    sc = yrtmi.Scanner(*a, maxRingDiff=mrd)
    ysc = Y.Scanner(*a, 1, mrd, 1, 1)
    im = yrtmi.ImageParams(*imdims)
    yim = Y.ImageParams.fromParams(*imdims)
    osc = oracle_scanner(sc)
    pos, ori = O.lut(osc)
    zb, ph, nr = O.histo_dims(osc)
    bins = np.arange(zb * ph * nr)
    d1, d2 = O.histo_pairs(osc, bins)
    L = O.lors_from_pairs(pos, ori, d1, d2)
    return sc, ysc, im, yim, osc, pos, ori, L


def test_histogram_frames_dd_psf_one_launch(gpu):
    """Three dynamic frames of a DD + projection-PSF histogram OSEM in one
    launch per subset (frame per bin), each frame against the oracle's
    reconstruction of its own histogram."""
    F, S, iters = 3, 3, 2
    sc, ysc, im, yim, osc, pos, ori, L = histo_setup(
        ("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12), 7, (20, 20, 12, 5.0, 5.0, 4.0))
    oi = oracle_image(im)
    rng = np.random.default_rng(23)
    k, s_step, ks = synth.gaussian_psf_row(2.0, k_spacing=0.25, num_s=2, s_step=50.0)
    po = O.proj_for_scanner(osc, type="dd", psf_kernels=k, psf_s_step=s_step,
                            psf_k_spacing=ks)
    ys = []
    for f in range(F):
        act = synth.cylinder_spheres(im, 40.0, 36.0, 1.0,
                                     [(5.0 - 4 * f, -4.0, 2.0, 8.0, 2.0 + 2.0 * f)],
                                     supersample=2)
        y = O.forward(oi, act[None], po, L)
        ys.append(rng.poisson(np.clip(y, 0, None) * (0.3 + 0.2 * f)).astype(np.float32))
    osem = Y.createOSEM(ysc)
    ctx = osem.getContext()
    osem.setImageParams(yim)
    osem.setProjector("DD")
    osem.getProjectorParams().setProjPsfKernels(k, s_step, ks)
    osem.num_OSEM_subsets = S
    osem.num_MLEM_iterations = iters
    hs = []
    for f in range(F):
        h = Y.Histogram3DDeviceOwned(ctx)
        h.copyFromHost(ys[f])
        hs.append(h)
    osem.setDataInput(hs)
    assert osem.numHistogramFrames() == F
    sens = [s_.transferToHostMemory() for s_ in osem.generateSensitivityImages()]
    x = osem.reconstruct().transferToHostMemory().reshape(F, im.nz, im.ny, im.nx)
    assert osem.numPlannedSubsets() == S
    osens = O.sensitivity_images(osc, oi, po, S, False, pos_ori=(pos, ori))
    for s in range(S):
        check(f"frames_sens_{s}", sens[s], osens[s], TOL_SENS)
    subsets = [O.histo_subset_bins(osc, S, s) for s in range(S)]
    for f in range(F):
        xo = O.reconstruct(osc, oi, po, L, osens, iters, S, False, meas=ys[f],
                           subset_indices=subsets)
        check(f"frames_dd_psf_frame{f}", x[f], xo, TOL_IMG)


def test_histogram_configuration_full_size(gpu):
    """The 24 x 256 histogram configuration at its full size: 18.9 M bins,
    96x96x48 @ 2 mm, water cylinder with 4 hot spheres, mu 0.096 /cm,
    5 iterations x 8 subsets, against the oracle flow (16 threads)."""
    S, iters = 8, 5
    a = ("cfg1", 96.0, 4.0, 4.0, 20.0, 300.0, 256, 24)
    sc, ysc, im, yim, osc, pos, ori, L = histo_setup(a, 23, (96, 96, 48, 2.0, 2.0, 2.0))
    oi = oracle_image(im)
    rng = np.random.default_rng(2024)
    spheres = synth.random_spheres_in_cylinder(rng, 100.0, 90.0, [5.0, 8.0, 11.0, 15.0], 4.0)
    act = synth.cylinder_spheres(im, 100.0, 90.0, 1.0, spheres, supersample=2)
    body = synth.cylinder_spheres(im, 100.0, 90.0, 1.0, [], supersample=1)
    mu = np.where(body > 0, np.float32(0.096), np.float32(0.0)).astype(np.float32)
    po = O.proj_for_scanner(osc)
    nt = 16
    acf = O.attenuation_factors(oi, mu, L, nt)
    y = O.forward(oi, act[None], po, L, nthreads=nt) * acf
    y = rng.poisson(np.clip(y, 0, None) * (50e6 / float(y.sum()))).astype(np.float32)
    osem = Y.createOSEM(ysc)
    ctx = osem.getContext()
    osem.setImageParams(yim)
    mu_img = Y.ImageDeviceOwned(ctx, yim)
    mu_img.copyFromHost(mu)
    osem.setAttenuationImage(mu_img)
    osem.num_OSEM_subsets = S
    osem.num_MLEM_iterations = iters
    h = Y.Histogram3DDeviceOwned(ctx)
    h.copyFromHost(y)
    osem.setDataInput(h)
    sens = [s_.transferToHostMemory() for s_ in osem.generateSensitivityImages()]
    x = osem.reconstruct().transferToHostMemory()
    osens = O.sensitivity_images(osc, oi, po, S, False, att_im=oi, att_img=mu,
                                 nthreads=nt, pos_ori=(pos, ori))
    for s in range(S):
        check(f"cfg1_sens_{s}", sens[s], osens[s], TOL_SENS)
    subsets = [O.histo_subset_bins(osc, S, s) for s in range(S)]
    xo = O.reconstruct(osc, oi, po, L, osens, iters, S, False, meas=y, att=acf,
                       subset_indices=subsets, nthreads=nt)
    check("cfg1_full_size", x, xo, TOL_IMG)


def listmode_osem(ysc, yim, ev, S, iters, out=None, **opts):
    osem = Y.createOSEM(ysc)
    ctx = osem.getContext()
    lm = Y.ListModeDeviceOwned(ctx, ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32),
                               tof=ev["tof"].astype(np.float32))
    osem.setImageParams(yim)
    osem.setNumRays(4)
    osem.addTOF(250.0, 5)
    osem.num_OSEM_subsets = S
    osem.num_MLEM_iterations = iters
    for k_, v in opts.items():
        setattr(osem, k_, v)
    osem.setDataInput(lm)
    osem.generateSensitivityImages()
    x = osem.reconstruct().transferToHostMemory()
    if out is not None:
        out["iteration_s"] = list(osem.iterationSeconds())
    return x, osem.numPlannedSubsets()


def test_memory_budget_graphs_and_autotune(gpu):
    """A plan budget below all the subsets' plans: the subsets beyond it run
    the unplanned EM kernel, the image does not change; graph-replayed
    iterations and autotuned launch shapes give the call-by-call defaults'
    image; all against the oracle."""
    S, iters, n = 4, 3, 60000
    a = ("lm", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    sc, ysc = yrtmi.Scanner(*a), Y.Scanner(*a, 1, 15, 1, 1)
    im = yrtmi.ImageParams(24, 24, 14, 6.0, 6.0, 4.5)
    yim = Y.ImageParams.fromParams(24, 24, 14, 6.0, 6.0, 4.5)
    act = synth.cylinder_spheres(im, 0.6 * im.nx * im.vx / 2, 0.7 * im.nz * im.vz, 1.0,
                                 [(5.0, -4.0, 2.0, 8.0, 4.0)], supersample=2)
    ev = synth.sample_events(sc, im, act, n, seed=5, tof_fwhm_ps=250.0)
    osc = oracle_scanner(sc)
    pos, ori = O.lut(osc)
    po = O.proj_for_scanner(osc, num_rays=4, tof_width_ps=250.0, tof_n_std=5)
    oim = oracle_image(im)
    osens = O.sensitivity_images(osc, oim, po, S, True, pos_ori=(pos, ori))
    L = O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev["tof"])
    xo = O.reconstruct(osc, oim, po, L, osens, iters, S, True)
    x_def, np_def = listmode_osem(ysc, yim, ev, S, iters)
    assert np_def == S
    check("budget_default", x_def, xo, TOL_IMG)
    x_plain, _ = listmode_osem(ysc, yim, ev, S, iters, useGraphs=False, useAutotune=False)
    check("no_graph_no_tune", x_plain, xo, TOL_IMG)
    check("graphs_autotune_vs_calls", x_def, x_plain, 1e-5)
    # iterations timed by OSEM_MI itself: one time per iteration, the image
    # as without timing
    timed = {}
    x_timed, _ = listmode_osem(ysc, yim, ev, S, iters, out=timed, timeIterations=True)
    print(f"iteration seconds {timed['iteration_s']}", flush=True)
    assert len(timed["iteration_s"]) == iters and min(timed["iteration_s"]) > 0
    check("timed_vs_untimed", x_timed, x_def, 1e-5)
    # a budget of one byte: no plan fits, every subset runs unplanned
    x_none, np_none = listmode_osem(ysc, yim, ev, S, iters, memoryBudget=1)
    assert np_none == 0
    check("budget_none", x_none, xo, TOL_IMG)
    # about two subsets' plans: the first subsets planned, the rest not
    ctx = yrtmi.Context(sc)
    pp = yrtmi.ProjectorParams(sc, numRays=4)
    pp.addTOF(250.0, 5)
    proj = ctx.projector(pp)
    import torch
    idx = np.arange(0, n, S)
    L0 = yrtmi.LorSet.pairs(torch.from_numpy(ev["d1"][idx].astype(np.int32)).to(gpu),
                            torch.from_numpy(ev["d2"][idx].astype(np.int32)).to(gpu),
                            tof=torch.from_numpy(ev["tof"][idx]).to(gpu))
    one = ctx.plan(proj, im, L0, tile_shape=yrtmi.capi.TILE_32x32x8).info()["bytes"]
    x_part, np_part = listmode_osem(ysc, yim, ev, S, iters, memoryBudget=int(2.5 * one),
                                    useAutotune=False)
    print(f"one subset's plan {one} B: {np_part} of {S} subsets planned", flush=True)
    assert 0 < np_part < S
    check("budget_partial", x_part, xo, TOL_IMG)


def test_histo_pair_bins_match_oracle(gpu):
    """Detector pair -> Histogram3D bin on the device (the reference's ring
    map, in-ring pairs of one position included) against the oracle's
    getBinIdFromDetPair restatement, and back to the pair."""
    import torch
    a = ("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12)
    sc = yrtmi.Scanner(*a, maxRingDiff=7)
    osc = oracle_scanner(sc)
    ctx = yrtmi.Context(sc)
    rng = np.random.default_rng(4)
    nd = sc.getNumDets()
    d1 = rng.integers(0, nd, 50000).astype(np.uint32)
    d2 = rng.integers(0, nd, 50000).astype(np.uint32)
    # a third of the pairs in one in-ring position on two rings
    d2[::3] = (d1[::3] % 64) + 64 * rng.integers(0, 12, d1[::3].size).astype(np.uint32)
    want = O.histo_bins_from_pairs(osc, d1, d2)
    ok = (want >= 0) & ((d1 != 0) | (d2 != 0))
    d1, d2, want = d1[ok], d2[ok], want[ok]
    assert (d1 % 64 == d2 % 64).sum() > 1000
    T = lambda v: torch.from_numpy(v.astype(np.int32)).to(gpu)
    got = ctx.histo_pair_bins(T(d1), T(d2)).cpu().numpy().astype(np.int64)
    np.testing.assert_array_equal(got, want)
    # back to pairs: the same unordered pair, except that a bin whose in-ring
    # pair has one position decodes to the disabled pair (0, 0)
    # (Histogram3D::getDetPairFromCoords, Histogram3D.cpp:317-323)
    b1, b2 = O.histo_pairs(osc, got)
    same = d1 % 64 == d2 % 64
    assert ((np.minimum(b1, b2) == np.minimum(d1, d2)) &
            (np.maximum(b1, b2) == np.maximum(d1, d2)))[~same].all()
    assert ((b1 == 0) & (b2 == 0))[same].all()
    # a pair beyond the ring-difference limit fails the call
    with pytest.raises(yrtmi.capi.YrtmiError, match="outside the histogram"):
        ctx.histo_pair_bins(T(np.array([1], np.uint32)), T(np.array([64 * 11 + 40], np.uint32)))
