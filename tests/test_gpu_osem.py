"""OSEM on the GPU (yrtmi.osem, through the C ABI) against the oracle's
restatement of the reference CPU OSEM flow, at small sizes of the benchmark
configurations.

Bars: sensitivity images <= 1e-5 relative L2; images after the iteration
count <= 1e-4 relative L2 (atomic-order nondeterminism included).
"""
import numpy as np
import pytest
import torch

import oracle as O
import yrtmi
from helpers import oracle_image, oracle_proj, oracle_scanner, rel_l2
from yrtmi import osem, synth

pytestmark = pytest.mark.gpu

TOL_SENS = 1e-5
TOL_IMG = 1e-4


def check(name, got, ref, tol):
    r = rel_l2(got, ref)
    print(f"[{name}] rel L2 = {r:.3e} (bar {tol:.0e})", flush=True)
    assert r <= tol, f"{name}: {r:.3e} > {tol:.0e}"


def hist_scanner():
    return yrtmi.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, maxRingDiff=7)


def lm_scanner():
    return yrtmi.Scanner("lm", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)


def phantom(im, rng):
    act = synth.cylinder_spheres(im, 0.6 * im.nx * im.vx / 2, 0.7 * im.nz * im.vz, 1.0,
                                 [(5.0, -4.0, 2.0, 8.0, 4.0)], supersample=2)
    return act


def run_histogram(gpu, projector, rays, psf, att, S, iters):
    sc = hist_scanner()
    im = yrtmi.ImageParams(20, 20, 12, 5.0, 5.0, 4.0)
    rng = np.random.default_rng(7)
    osc = oracle_scanner(sc)
    pos, ori = O.lut(osc)
    zb, ph, nr = O.histo_dims(osc)
    bins = np.arange(zb * ph * nr)
    d1, d2 = O.histo_pairs(osc, bins)
    L = O.lors_from_pairs(pos, ori, d1, d2)
    act = phantom(im, rng)
    mu = None
    if att:
        mu = np.where(act > 0, np.float32(0.096), np.float32(0.0)).astype(np.float32)
    po_true = O.proj_for_scanner(osc, type=projector.lower(), num_rays=rays)
    y = O.forward(oracle_image(im), act[None], po_true, L)
    if mu is not None:
        y = y * O.attenuation_factors(oracle_image(im), mu, L)
    y = rng.poisson(np.clip(y, 0, None) * 0.5).astype(np.float32)
    # GPU
    rec = osem.createOSEM(sc)
    rec.setImageParams(im)
    rec.setProjector(projector)
    rec.setNumRays(rays)
    if psf:
        k, s_step, ks = synth.gaussian_psf_row(2.0, k_spacing=0.25, num_s=2, s_step=50.0)
        rec.addProjPSF(k, s_step, ks)
    if att:
        rec.setAttenuationImage(torch.from_numpy(mu).to(gpu), im)
    rec.num_OSEM_subsets = S
    rec.num_MLEM_iterations = iters
    rec.setDataInput(osem.Histogram3DData(torch.from_numpy(y.reshape(zb, ph, nr)).to(gpu)))
    sens = rec.generateSensitivityImages()
    x = rec.reconstruct()
    torch.cuda.synchronize()
    # oracle
    pp = rec.projectorParams
    po = oracle_proj(pp)
    osens = O.sensitivity_images(osc, oracle_image(im), po, S, False,
                                 att_im=oracle_image(im) if att else None, att_img=mu,
                                 pos_ori=(pos, ori))
    for s in range(S):
        check(f"sens_{projector}_{s}", sens[s].cpu().numpy(), osens[s], TOL_SENS)
    acf = O.attenuation_factors(oracle_image(im), mu, L) if att else None
    subsets = [O.histo_subset_bins(osc, S, s) for s in range(S)]
    xo = O.reconstruct(osc, oracle_image(im), po, L, osens, iters, S, False, meas=y, att=acf,
                       subset_indices=subsets)
    check(f"osem_{projector}{rays}_psf{psf}_att{att}", x.cpu().numpy(), xo, TOL_IMG)


def test_osem_histogram_siddon_attenuation(gpu):
    run_histogram(gpu, "Siddon", 1, False, True, 4, 3)


def test_osem_histogram_dd_psf(gpu):
    run_histogram(gpu, "DD", 1, True, False, 3, 2)


def run_listmode(gpu, motion=False, frames=1, n=60000, S=5, iters=2):
    sc = lm_scanner()
    im = yrtmi.ImageParams(24, 24, 14, 6.0, 6.0, 4.5)
    rng = np.random.default_rng(11)
    act = phantom(im, rng)
    ev = synth.sample_events(sc, im, act, n, seed=5, tof_fwhm_ps=250.0)
    xf = None
    if motion:
        xf = synth.rigid_transforms(rng, 4)
        ev["motion"] = rng.integers(0, 4, n).astype(np.uint8)
    if frames > 1:
        ev["frame"] = rng.integers(0, frames, n).astype(np.int32)
    T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(gpu)
    data = osem.ListModeData(T(ev["d1"], np.int32), T(ev["d2"], np.int32), tof=T(ev["tof"], np.float32),
                             frame=T(ev["frame"], np.int32) if frames > 1 else None, num_frames=frames,
                             motion_frame=T(ev["motion"], np.uint8) if motion else None,
                             motion_transforms=xf)
    rec = osem.createOSEM(sc)
    rec.setImageParams(im)
    rec.setNumRays(4)
    rec.addTOF(250.0, 5)
    rec.num_OSEM_subsets = S
    rec.num_MLEM_iterations = iters
    rec.setDataInput(data)
    sens = rec.generateSensitivityImages()
    x = rec.reconstruct()
    torch.cuda.synchronize()
    osc = oracle_scanner(sc)
    pos, ori = O.lut(osc)
    po = oracle_proj(rec.projectorParams)
    oim = oracle_image(im)
    osens = O.sensitivity_images(osc, oim, po, S, True, pos_ori=(pos, ori))
    check("sens_lm", sens[0].cpu().numpy(), osens[0], TOL_SENS)
    L = O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev["tof"], frame=ev.get("frame"),
                          motion_frame=ev.get("motion"), motion_xf=xf)
    oim_t = O.ImageDef(im.nx, im.ny, im.nz, im.vx, im.vy, im.vz, nt=frames)
    xo = O.reconstruct(osc, oim_t, po, L, osens, iters, S, True)
    check(f"osem_lm_motion{motion}_frames{frames}", x.cpu().numpy(), xo, TOL_IMG)


def test_osem_listmode_tof_multiray(gpu):
    run_listmode(gpu)


def test_osem_listmode_motion(gpu):
    run_listmode(gpu, motion=True)


def test_osem_listmode_dynamic_frames(gpu):
    run_listmode(gpu, frames=3, n=90000, S=3)
