"""The attenuation split through the C++ host layer (OSEM_MI's hardware /
in-vivo attenuation images and total / hardware / in-vivo ACF histograms,
resolved as Corrector::setup does, Corrector.cpp:34-166) against the
oracle's flow: sensitivity images weighted by the hardware attenuation, the
EM model by the total attenuation, the EM denominator precorrected by the
in-vivo attenuation (OSEM_GPU.cu:1048-1110, 1159-1341).  Bars: 1e-5 per
sensitivity image and 1e-4 per reconstruction, relative L2."""
import numpy as np
import pytest

import oracle as O
import pyyrtmi as Y
import yrtmi
from helpers import oracle_image, oracle_scanner, rel_l2
from yrtmi import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["images", "acf_split", "acf_total"])
def test_osem_histogram_attenuation_split(gpu, mode):
    sc = yrtmi.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, maxRingDiff=7)
    ysc = Y.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, 1, 7, 1, 1)
    im = yrtmi.ImageParams(40, 40, 20, 3.0, 3.0, 2.5)
    yim = Y.ImageParams.fromParams(40, 40, 20, 3.0, 3.0, 2.5)
    osc, oi = oracle_scanner(sc), oracle_image(im)
    pos, ori = O.lut(osc)
    zb, ph, nr = O.histo_dims(osc)
    bins = np.arange(zb * ph * nr)
    d1, d2 = O.histo_pairs(osc, bins)
    L = O.lors_from_pairs(pos, ori, d1, d2)
    po = O.proj_for_scanner(osc)
    act = synth.cylinder_spheres(im, 40.0, 30.0, 1.0, [(5.0, 3.0, 0.0, 8.0, 4.0)],
                                 supersample=1)
    # in-vivo: water in the body; hardware: a dense bed slab under it
    iv = synth.cylinder_spheres(im, 45.0, 40.0, 0.096, [], supersample=1)[None].astype(np.float32)
    hw = np.zeros_like(iv)
    hw[..., 8:11, :] = 0.3
    acf_iv = O.attenuation_factors(oi, iv, L)
    acf_hw = O.attenuation_factors(oi, hw, L)
    acf_tot = O.attenuation_factors(oi, (hw + iv).astype(np.float32), L)
    rng = np.random.default_rng(21)
    y = rng.poisson(np.clip(O.forward(oi, act[None], po, L), 0, None) * acf_tot * 0.5
                    ).astype(np.float32)
    S, iters = 2, 2
    osem = Y.createOSEM(ysc)
    ctx = osem.getContext()
    osem.setImageParams(yim)
    osem.num_OSEM_subsets = S
    osem.num_MLEM_iterations = iters
    h = Y.Histogram3DDeviceOwned(ctx)
    h.copyFromHost(y)
    osem.setDataInput(h)
    keep = []

    def image(a):
        d = Y.ImageDeviceOwned(ctx, yim)
        d.copyFromHost(a)
        keep.append(d)
        return d

    def histo(a):
        d = Y.Histogram3DDeviceOwned(ctx)
        d.copyFromHost(np.ascontiguousarray(a, np.float32))
        keep.append(d)
        return d

    if mode == "images":
        osem.setHardwareAttenuationImage(image(hw))
        osem.setInVivoAttenuationImage(image(iv))
        sens_kw = dict(att_im=oi, att_img=hw)
        att, pre = acf_tot, acf_iv
    elif mode == "acf_split":
        osem.setHardwareACFHistogram(histo(acf_hw))
        osem.setInVivoACFHistogram(histo(acf_iv))
        sens_kw = dict(acf_histo=acf_hw)
        att, pre = (acf_iv * acf_hw).astype(np.float32), acf_iv
    else:
        # a total ACF alone is all hardware: no precorrection
        osem.setACFHistogram(histo(acf_tot))
        sens_kw = dict(acf_histo=acf_tot)
        att, pre = acf_tot, None
    sens = [s_.transferToHostMemory() for s_ in osem.generateSensitivityImages()]
    osens = O.sensitivity_images(osc, oi, po, S, False, pos_ori=(pos, ori), **sens_kw)
    for s in range(S):
        r = rel_l2(sens[s], osens[s])
        print(f"{mode} sens {s}: {r:.2e}", flush=True)
        assert r <= 1e-5
    x = osem.reconstruct().transferToHostMemory()
    subsets = [O.histo_subset_bins(osc, S, s) for s in range(S)]
    xo = O.reconstruct(osc, oi, po, L, osens, iters, S, False, meas=y, subset_indices=subsets,
                       att=att, precorr=pre)
    r = rel_l2(x, xo)
    print(f"{mode} osem: {r:.2e}", flush=True)
    assert r <= 1e-4
    if pre is not None:
        # the precorrection is applied
        xn = O.reconstruct(osc, oi, po, L, osens, iters, S, False, meas=y,
                           subset_indices=subsets, att=att)
        assert rel_l2(x, xn) > 1e-3


@pytest.mark.parametrize("mode", ["acf_split", "acf_total"])
def test_osem_listmode_acf_histograms(gpu, mode):
    """ACF histograms read at each event's detector pair (list-mode)."""
    sc = yrtmi.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, maxRingDiff=7)
    ysc = Y.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, 1, 7, 1, 1)
    im = yrtmi.ImageParams(40, 40, 20, 3.0, 3.0, 2.5)
    yim = Y.ImageParams.fromParams(40, 40, 20, 3.0, 3.0, 2.5)
    osc, oi = oracle_scanner(sc), oracle_image(im)
    pos, ori = O.lut(osc)
    zb, ph, nr = O.histo_dims(osc)
    hb = np.arange(zb * ph * nr)
    L = O.lors_from_pairs(pos, ori, *O.histo_pairs(osc, hb))
    po = O.proj_for_scanner(osc)
    iv = synth.cylinder_spheres(im, 45.0, 40.0, 0.096, [], supersample=1)[None].astype(np.float32)
    hw = np.zeros_like(iv)
    hw[..., 8:11, :] = 0.3
    acf_iv = O.attenuation_factors(oi, iv, L)
    acf_hw = O.attenuation_factors(oi, hw, L)
    act = synth.cylinder_spheres(im, 40.0, 30.0, 1.0, [(5.0, 3.0, 0.0, 8.0, 4.0)],
                                 supersample=1)
    ev = synth.sample_events(sc, im, act, 60000, seed=23)
    d1, d2 = ev["d1"].astype(np.uint32), ev["d2"].astype(np.uint32)
    ring = lambda d: (d // 64) % 12
    keep = np.abs(ring(d1).astype(int) - ring(d2).astype(int)) <= 7
    d1, d2 = d1[keep], d2[keep]
    bins = O.histo_bins_from_pairs(osc, d1, d2)
    assert np.all(bins >= 0)
    S, iters = 3, 2
    osem = Y.createOSEM(ysc)
    ctx = osem.getContext()
    osem.setImageParams(yim)
    osem.num_OSEM_subsets = S
    osem.num_MLEM_iterations = iters
    lm = Y.ListModeDeviceOwned(ctx, d1, d2)
    osem.setDataInput(lm)
    hs = []
    for a in ((acf_iv, acf_hw) if mode == "acf_split" else (acf_iv * acf_hw,)):
        h = Y.Histogram3DDeviceOwned(ctx)
        h.copyFromHost(np.ascontiguousarray(a, np.float32))
        hs.append(h)
    if mode == "acf_split":
        osem.setInVivoACFHistogram(hs[0])
        osem.setHardwareACFHistogram(hs[1])
        hw_h, att, pre = acf_hw, (acf_iv[bins] * acf_hw[bins]).astype(np.float32), acf_iv[bins]
    else:
        tot = (acf_iv * acf_hw).astype(np.float32)
        osem.setACFHistogram(hs[0])
        hw_h, att, pre = tot, tot[bins], None
    sens = osem.generateSensitivityImages()[0].transferToHostMemory()
    osens = O.sensitivity_images(osc, oi, po, S, True, pos_ori=(pos, ori), acf_histo=hw_h)
    r = rel_l2(sens, osens[0])
    print(f"lm {mode} sens: {r:.2e}", flush=True)
    assert r <= 1e-5
    x = osem.reconstruct().transferToHostMemory()
    Le = O.lors_from_pairs(pos, ori, d1, d2)
    xo = O.reconstruct(osc, oi, po, Le, osens, iters, S, True, att=att, precorr=pre)
    r = rel_l2(x, xo)
    print(f"lm {mode} osem: {r:.2e}", flush=True)
    assert r <= 1e-4


def test_listmode_acf_rejects_pairs_outside_the_histogram(gpu):
    ysc = Y.Scanner("h", 48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, 1, 7, 1, 1)
    osem = Y.createOSEM(ysc)
    ctx = osem.getContext()
    osem.setImageParams(Y.ImageParams.fromParams(20, 20, 10, 4.0, 4.0, 4.0))
    # ring 0 to ring 10: a ring difference above the limit of 7
    lm = Y.ListModeDeviceOwned(ctx, np.array([5, 3], np.uint32),
                               np.array([64 * 10 + 37, 64 + 35], np.uint32))
    osem.setDataInput(lm)
    zb, ph, nr = O.histo_dims(O.ScannerDef(48.0, 4.0, 4.0, 10.0, 110.0, 64, 12,
                                           max_ring_diff=7))
    h = Y.Histogram3DDeviceOwned(ctx)
    h.copyFromHost(np.ones(zb * ph * nr, np.float32))
    osem.setInVivoACFHistogram(h)
    osem.generateSensitivityImages()
    with pytest.raises(Exception, match="outside the histogram"):
        osem.reconstruct()
