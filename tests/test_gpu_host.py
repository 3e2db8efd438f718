"""The C++ host layer (pyyrtmi: OSEM, OperatorProjectorMI, forwProject /
backProject with the reference's names) on the GPU against the oracle's
restatement of the reference CPU path, at small sizes.

Bars: projections and backprojections <= 1e-5 relative L2; sensitivity
images <= 1e-5; images after the iteration count <= 1e-4 (atomic-order
nondeterminism included).
"""
import numpy as np
import pytest
import torch

import oracle as O
import pyyrtmi as Y
import yrtmi
from helpers import oracle_image, oracle_scanner, rel_l2
from yrtmi import synth

pytestmark = pytest.mark.gpu

TOL_PROJ = 1e-5
TOL_SENS = 1e-5
TOL_IMG = 1e-4


def check(name, got, ref, tol):
    r = rel_l2(got, ref)
    print(f"[{name}] rel L2 = {r:.3e} (bar {tol:.0e})", flush=True)
    assert r <= tol, f"{name}: {r:.3e} > {tol:.0e}"


def scanners(kind):
    """The same scanner as the Python geometry class (data generation, oracle)
    and as the host layer's class."""
    if kind == "h":
        a = ("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12)
        return yrtmi.Scanner(*a, maxRingDiff=7), Y.Scanner(*a, 1, 7, 1, 1)
    a = ("lm", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    return yrtmi.Scanner(*a), Y.Scanner(*a, 1, 15, 1, 1)


def images(nx, ny, nz, vx, vy, vz, nt=1):
    return (yrtmi.ImageParams(nx, ny, nz, vx, vy, vz, nt=nt),
            Y.ImageParams.fromParams(nx, ny, nz, vx, vy, vz, nt))


def phantom(im):
    return synth.cylinder_spheres(im, 0.6 * im.nx * im.vx / 2, 0.7 * im.nz * im.vz, 1.0,
                                  [(5.0, -4.0, 2.0, 8.0, 4.0)], supersample=2)


def run_histogram(gpu, projector, psf, att, S, iters, sens_histo=False, invert=False,
                  randoms=False):
    sc, ysc = scanners("h")
    osc = oracle_scanner(sc)
    im, yim = images(20, 20, 12, 5.0, 5.0, 4.0)
    oi = oracle_image(im)
    rng = np.random.default_rng(7)
    pos, ori = O.lut(osc)
    zb, ph, nr = O.histo_dims(osc)
    bins = np.arange(zb * ph * nr)
    d1, d2 = O.histo_pairs(osc, bins)
    L = O.lors_from_pairs(pos, ori, d1, d2)
    act = phantom(im)
    mu = np.where(act > 0, np.float32(0.096), np.float32(0.0)).astype(np.float32) if att else None
    y = O.forward(oi, act[None], O.proj_for_scanner(osc, type=projector.lower()), L)
    if att:
        y = y * O.attenuation_factors(oi, mu, L)
    y = rng.poisson(np.clip(y, 0, None) * 0.5).astype(np.float32)
    norm = rng.uniform(0.5, 1.5, y.size).astype(np.float32) if sens_histo else None
    if norm is not None:
        norm[rng.integers(0, y.size, 50)] = 0.0
    rnd = np.full(y.size, 0.05, np.float32) if randoms else None
    k = s_step = ks = None
    if psf:
        k, s_step, ks = synth.gaussian_psf_row(2.0, k_spacing=0.25, num_s=2, s_step=50.0)

    osem = Y.createOSEM(ysc)
    ctx = osem.getContext()
    osem.setImageParams(yim)
    osem.setProjector(projector)
    if psf:
        osem.getProjectorParams().setProjPsfKernels(k, s_step, ks)
    if att:
        mu_img = Y.ImageDeviceOwned(ctx, yim)
        mu_img.copyFromHost(mu)
        osem.setAttenuationImage(mu_img)
    if norm is not None:
        nh = Y.Histogram3DDeviceOwned(ctx)
        nh.copyFromHost(norm)
        osem.setSensitivityHistogram(nh)
        osem.setInvertSensitivity(invert)
    if rnd is not None:
        rh = Y.Histogram3DDeviceOwned(ctx)
        rh.copyFromHost(rnd)
        osem.setRandomsHistogram(rh)
    osem.num_OSEM_subsets = S
    osem.num_MLEM_iterations = iters
    h = Y.Histogram3DDeviceOwned(ctx)
    h.copyFromHost(y)
    osem.setDataInput(h)
    sens = [s_.transferToHostMemory() for s_ in osem.generateSensitivityImages()]
    x = osem.reconstruct().transferToHostMemory()
    # oracle
    po = O.proj_for_scanner(osc, type=projector.lower(), psf_kernels=k,
                            psf_s_step=s_step or 0.0, psf_k_spacing=ks or 0.0)
    osens = O.sensitivity_images(osc, oi, po, S, False, att_im=oi if att else None,
                                 att_img=mu, pos_ori=(pos, ori), sens_histo=norm,
                                 invert_sens=invert)
    tag = f"{projector}_psf{psf}_att{att}_norm{sens_histo}{'inv' if invert else ''}"
    for s in range(S):
        check(f"host_sens_{tag}_{s}", sens[s], osens[s], TOL_SENS)
    acf = O.attenuation_factors(oi, mu, L) if att else None
    sf = O.sensitivity_factors(norm, bins, invert) if norm is not None else None
    subsets = [O.histo_subset_bins(osc, S, s) for s in range(S)]
    xo = O.reconstruct(osc, oi, po, L, osens, iters, S, False, meas=y, att=acf, rnd=rnd,
                       subset_indices=subsets, norm=sf)
    check(f"host_osem_{tag}", x, xo, TOL_IMG)


def test_host_osem_histogram_siddon_attenuation(gpu):
    run_histogram(gpu, "Siddon", False, True, 4, 3)


def test_host_osem_histogram_dd_psf(gpu):
    run_histogram(gpu, "DD", True, False, 3, 2)


def test_host_osem_histogram_normalisation_randoms(gpu):
    run_histogram(gpu, "Siddon", False, False, 2, 2, sens_histo=True, invert=True,
                  randoms=True)


def run_listmode(gpu, motion=False, frames=1, n=60000, S=5, iters=2, tiles=True):
    sc, ysc = scanners("lm")
    osc = oracle_scanner(sc)
    im, yim = images(24, 24, 14, 6.0, 6.0, 4.5)
    rng = np.random.default_rng(11)
    ev = synth.sample_events(sc, im, phantom(im), n, seed=5, tof_fwhm_ps=250.0)
    xf = None
    if motion:
        xf = synth.rigid_transforms(rng, 4)
        ev["motion"] = rng.integers(0, 4, n).astype(np.uint8)
    if frames > 1:
        ev["frame"] = rng.integers(0, frames, n).astype(np.int32)
    osem = Y.createOSEM(ysc)
    ctx = osem.getContext()
    lm = Y.ListModeDeviceOwned(ctx, ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32),
                               tof=ev["tof"].astype(np.float32),
                               frame=ev.get("frame"), motion=ev.get("motion"))
    lm.numFrames = frames
    if xf is not None:
        lm.setMotionTransforms(xf)
    osem.setImageParams(yim)
    osem.setNumRays(4)
    osem.addTOF(250.0, 5)
    osem.useTilePlans = tiles
    osem.num_OSEM_subsets = S
    osem.num_MLEM_iterations = iters
    osem.setDataInput(lm)
    sens = osem.generateSensitivityImages()[0].transferToHostMemory()
    x = osem.reconstruct().transferToHostMemory()
    pos, ori = O.lut(osc)
    po = O.proj_for_scanner(osc, num_rays=4, tof_width_ps=250.0, tof_n_std=5)
    osens = O.sensitivity_images(osc, oracle_image(im), po, S, True, pos_ori=(pos, ori))
    check("host_sens_lm", sens, osens[0], TOL_SENS)
    L = O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev["tof"], frame=ev.get("frame"),
                          motion_frame=ev.get("motion"), motion_xf=xf)
    oim_t = O.ImageDef(im.nx, im.ny, im.nz, im.vx, im.vy, im.vz, nt=frames)
    xo = O.reconstruct(osc, oim_t, po, L, osens, iters, S, True)
    check(f"host_osem_lm_motion{motion}_frames{frames}_tiles{tiles}", x, xo, TOL_IMG)


@pytest.mark.parametrize("tiles", [True, False])
def test_host_osem_listmode_tof_multiray(gpu, tiles):
    run_listmode(gpu, tiles=tiles)


def test_host_osem_listmode_motion(gpu):
    run_listmode(gpu, motion=True)


def test_host_osem_listmode_dynamic_frames(gpu):
    run_listmode(gpu, frames=3, n=90000, S=3)


def test_host_operator_projector_and_utils(gpu):
    sc, ysc = scanners("lm")
    osc = oracle_scanner(sc)
    ctx = Y.Context(ysc)
    im, yim = images(24, 24, 14, 6.0, 6.0, 4.5)
    rng = np.random.default_rng(3)
    ev = synth.random_events(sc, 5000, rng, tof_fwhm_ps=250.0)
    lm = Y.ListModeDeviceOwned(ctx, ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32),
                               tof=ev["tof"].astype(np.float32))
    pp = Y.ProjectorParams(ysc)
    pp.numRays = 4
    pp.addTOF(250.0, 5)
    x = rng.uniform(0, 2, (im.nz, im.ny, im.nx)).astype(np.float32)
    img = Y.ImageDeviceOwned(ctx, yim)
    img.copyFromHost(x)
    op = Y.OperatorProjectorMI(ctx, pp)
    op.applyA(img, lm)
    ctx.synchronize()
    fwd = lm.getProjectionValues()
    pos, ori = O.lut(osc)
    L = O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev["tof"])
    po = O.proj_for_scanner(osc, num_rays=4, tof_width_ps=250.0, tof_n_std=5)
    oi = oracle_image(im)
    check("host_applyA", fwd, O.forward(oi, x[None], po, L), TOL_PROJ)
    vals = rng.uniform(0.1, 1.0, 5000).astype(np.float32)
    lm.setProjectionValues(vals)
    out = Y.ImageDeviceOwned(ctx, yim)
    Y.backProject(ctx, pp, out, lm)
    check("host_backProject", out.transferToHostMemory(), O.back(oi, po, L, vals), TOL_PROJ)
    # histogram forward through the util function (every bin, bin order)
    sch, ysch = scanners("h")
    ctxh = Y.Context(ysch)
    imh, yimh = images(20, 20, 12, 5.0, 5.0, 4.0)
    xh = rng.uniform(0, 2, (imh.nz, imh.ny, imh.nx)).astype(np.float32)
    ih = Y.ImageDeviceOwned(ctxh, yimh)
    ih.copyFromHost(xh)
    ph = Y.ProjectorParams(ysch)
    hh = Y.Histogram3DDeviceOwned(ctxh)
    Y.forwProject(ctxh, ph, ih, hh)
    osh = oracle_scanner(sch)
    zb, nphi, nr = O.histo_dims(osh)
    bins = np.arange(zb * nphi * nr)
    d1, d2 = O.histo_pairs(osh, bins)
    posh, orih = O.lut(osh)
    Lh = O.lors_from_pairs(posh, orih, d1, d2)
    check("host_forwProject_histogram", hh.transferToHostMemory(),
          O.forward(oracle_image(imh), xh[None], O.proj_for_scanner(osh), Lh), TOL_PROJ)
    assert ctxh.histogramShape() == (zb, nphi, nr)
    with pytest.raises(RuntimeError):
        Y.createOSEM(ysc, use_gpu=False)
