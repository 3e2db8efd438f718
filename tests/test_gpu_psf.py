"""Image-space PSF on the GPU (yrtmi_psf_apply through the C ABI; OSEM with
an image PSF through the C++ host layer) against the oracle's restatement
of OperatorPsf::convolve and of OSEM_CPU's PSF steps.

Bars: the convolution is bitwise the oracle's (same taps, same FP32
multiply-then-add order); the reconstruction with PSF <= 1e-4 relative L2
against the oracle flow, as every OSEM image test."""
import json

import numpy as np
import pytest
import torch

import oracle as O
import pyyrtmi as Y
import yrtmi
from helpers import oracle_image, oracle_scanner, rel_l2
from yrtmi import cli, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,ks,nt", [((20, 24, 28), (5, 3, 7), 1),
                                         ((9, 11, 13), (9, 5, 1), 2),
                                         ((33, 17, 65), (11, 9, 13), 1)])
def test_psf_matches_oracle_bitwise(gpu, shape, ks, nt):
    nz, ny, nx = shape
    im = yrtmi.ImageParams(nx, ny, nz, 2.0, 2.5, 3.0, nt=nt)
    ctx = yrtmi.Context(yrtmi.Scanner("t", 96.0, 4.0, 4.0, 10.0, 180.0, 128, 24))
    rng = np.random.default_rng(sum(shape))
    x = rng.random((nt, nz, ny, nx), dtype=np.float32)
    kx, ky, kz = (rng.random(n).astype(np.float32) for n in ks)
    xg = torch.from_numpy(x).to(gpu)
    out = torch.empty_like(xg)
    tmp = torch.empty_like(xg)
    for transpose in (False, True):
        ctx.psf(im, xg, out, tmp, kx, ky, kz, transpose=transpose)
        ctx.synchronize()
        want = O.psf_convolve(x, kx, ky, kz, transpose=transpose)
        np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_psf_rejects_bad_arguments(gpu):
    im = yrtmi.ImageParams(8, 8, 8, 2.0, 2.0, 2.0)
    ctx = yrtmi.Context(yrtmi.Scanner("t", 96.0, 4.0, 4.0, 10.0, 180.0, 128, 24))
    x = torch.zeros(im.shape, device=gpu)
    o = torch.zeros_like(x)
    t = torch.zeros_like(x)
    k3 = np.ones(3, np.float32) / 3
    with pytest.raises(yrtmi.capi.YrtmiError, match="odd"):
        ctx.psf(im, x, o, t, np.ones(4, np.float32), k3, k3)
    with pytest.raises(yrtmi.capi.YrtmiError, match="half length"):
        ctx.psf(im, x, o, t, np.ones(17, np.float32), k3, k3)
    with pytest.raises(yrtmi.capi.YrtmiError, match="distinct"):
        ctx.psf(im, x, x, t, k3, k3, k3)


def test_osem_histogram_with_image_psf(gpu, tmp_path):
    """Histogram OSEM (Siddon) with a file-given image PSF: sensitivity
    images through PSF^T, forward through PSF, update through PSF^T."""
    sc = yrtmi.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, maxRingDiff=7)
    ysc = Y.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, 1, 7, 1, 1)
    im = yrtmi.ImageParams(40, 40, 20, 3.0, 3.0, 2.5)
    yim = Y.ImageParams.fromParams(40, 40, 20, 3.0, 3.0, 2.5)
    osc = oracle_scanner(sc)
    oi = oracle_image(im)
    pos, ori = O.lut(osc)
    zb, ph, nr = O.histo_dims(osc)
    bins = np.arange(zb * ph * nr)
    d1, d2 = O.histo_pairs(osc, bins)
    L = O.lors_from_pairs(pos, ori, d1, d2)
    act = synth.cylinder_spheres(im, 40.0, 30.0, 1.0, [(5.0, 3.0, 0.0, 8.0, 4.0)],
                                 supersample=1)
    po = O.proj_for_scanner(osc)
    rng = np.random.default_rng(12)
    y = rng.poisson(np.clip(O.forward(oi, act[None], po, L), 0, None) * 0.5).astype(np.float32)
    k = (O.gaussian_kernel_1d(1.5, 3.0, 5), O.gaussian_kernel_1d(1.5, 3.0, 3),
         O.gaussian_kernel_1d(2.0, 2.5, 5))
    f = str(tmp_path / "psf.csv")
    with open(f, "w") as fh:
        for kk in k:
            fh.write(",".join(repr(float(v)) for v in kk) + "\n")
        fh.write(",".join(str(len(kk)) for kk in k) + "\n")
    S, iters = 2, 2
    osem = Y.createOSEM(ysc)
    ctx = osem.getContext()
    osem.setImageParams(yim)
    osem.addImagePSF(f)
    assert osem.hasImagePSF()
    for a, b in zip(osem.getImagePSFKernels(), k):
        np.testing.assert_array_equal(np.float32(a), b)
    osem.num_OSEM_subsets = S
    osem.num_MLEM_iterations = iters
    h = Y.Histogram3DDeviceOwned(ctx)
    h.copyFromHost(y)
    osem.setDataInput(h)
    sens = [s_.transferToHostMemory() for s_ in osem.generateSensitivityImages()]
    x = osem.reconstruct().transferToHostMemory()
    osens = O.sensitivity_images(osc, oi, po, S, False, pos_ori=(pos, ori), image_psf=k)
    for s in range(S):
        r = rel_l2(sens[s], osens[s])
        print(f"psf sens {s}: {r:.2e}", flush=True)
        assert r <= 1e-5
    subsets = [O.histo_subset_bins(osc, S, s) for s in range(S)]
    xo = O.reconstruct(osc, oi, po, L, osens, iters, S, False, meas=y, subset_indices=subsets,
                       image_psf=k)
    r = rel_l2(x, xo)
    print(f"psf osem: {r:.2e}", flush=True)
    assert r <= 1e-4
    # the same without the PSF differs (the PSF is applied)
    xo0 = O.reconstruct(osc, oi, po, L, O.sensitivity_images(osc, oi, po, S, False,
                                                              pos_ori=(pos, ori)),
                        iters, S, False, meas=y, subset_indices=subsets)
    assert rel_l2(x, xo0) > 1e-2


def test_gaussian_image_psf_kernels(gpu):
    ysc = Y.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, 1, 7, 1, 1)
    osem = Y.createOSEM(ysc)
    osem.setImageParams(Y.ImageParams.fromParams(40, 40, 20, 3.0, 3.0, 2.5))
    osem.addUniformGaussianImagePSFFromSigma(2.0, 2.5, 3.0)
    got = osem.getImagePSFKernels()
    for g, (s, v) in zip(got, ((2.0, 3.0), (2.5, 3.0), (3.0, 2.5))):
        np.testing.assert_allclose(np.float32(g), O.gaussian_kernel_1d(s, v), rtol=1e-6)
    fwhm = 2.0 * np.sqrt(2.0 * np.log(2.0))
    osem.addUniformGaussianImagePSFFromFWHM(4.0, 4.0, 4.0, 7, 5, 3)
    got = osem.getImagePSFKernels()
    assert [len(g) for g in got] == [7, 5, 3]
    np.testing.assert_allclose(np.float32(got[0]), O.gaussian_kernel_1d(4.0 / fwhm, 3.0, 7),
                               rtol=1e-6)


def test_cli_psf(gpu, tmp_path):
    """--psf: forward_project blurs the image first, backproject applies the
    adjoint after (ForwardProject.cpp, Backproject.cpp)."""
    sj = str(tmp_path / "s.json")
    with open(sj, "w") as fh:
        json.dump({"VERSION": 3.2, "scannerName": "p", "axialFOV": 48.0,
                   "crystalSize_trans": 4.0, "crystalSize_z": 4.0, "crystalDepth": 10.0,
                   "scannerRadius": 110.0, "dets_per_ring": 64, "num_rings": 12,
                   "num_doi": 1, "max_ring_diff": 11, "min_ang_diff": 1}, fh)
    pj = str(tmp_path / "p.json")
    with open(pj, "w") as fh:
        json.dump({"VERSION": 1.2, "nx": 30, "ny": 30, "nz": 16, "vx": 4.0, "vy": 4.0,
                   "vz": 3.0}, fh)
    pf = str(tmp_path / "psf.csv")
    k = [O.gaussian_kernel_1d(2.0, 4.0, 5), O.gaussian_kernel_1d(2.0, 4.0, 5),
         O.gaussian_kernel_1d(2.0, 3.0, 3)]
    with open(pf, "w") as fh:
        for kk in k:
            fh.write(" ".join(repr(float(v)) for v in kk) + "\n")
        fh.write("5 5 3\n")
    sc = Y.Scanner(sj)
    ctx = Y.Context(sc)
    params = Y.ImageParams(pj)
    rng = np.random.default_rng(13)
    x = rng.random((1, 16, 30, 30), dtype=np.float32)
    xin = str(tmp_path / "x.nii")
    Y.writeImage(xin, params, x)
    hf = str(tmp_path / "h.rawd")
    assert cli.main(["forward_project", "-s", sj, "-i", xin, "-o", hf, "--psf", pf]) == 0
    img = Y.ImageDeviceOwned(ctx, params)
    img.copyFromHost(O.psf_convolve(x, *k))
    h = Y.Histogram3DDeviceOwned(ctx)
    Y.forwProject(ctx, Y.ProjectorParams(sc), img, h)
    ctx.synchronize()
    assert rel_l2(Y.readRawd(hf).reshape(-1), h.transferToHostMemory()) <= 1e-6
    bf = str(tmp_path / "b.nii")
    assert cli.main(["backproject", "-s", sj, "-i", hf, "-f", "H", "-p", pj, "-o", bf,
                     "--psf", pf]) == 0
    _, b_cli = Y.readImage(bf)
    b = Y.ImageDeviceOwned(ctx, params)
    b.setValue(0.0)
    Y.backProject(ctx, Y.ProjectorParams(sc), b, h)
    ctx.synchronize()
    want = O.psf_convolve(b.transferToHostMemory(), *k, transpose=True)
    assert rel_l2(b_cli, want) <= 1e-6
