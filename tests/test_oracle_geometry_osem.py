"""Pins of the oracle's data model (LUT, Histogram3D, subset iterators) and
OSEM pieces (EM update, threshold, EM subset) by closed forms and MLEM
invariants; the restated reference unit tests are cited per test."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))
FS = KA["fake_scanner"]
FAKE = O.ScannerDef(FS["axial_fov"], FS["crystal_size_z"], FS["crystal_size_trans"],
                    FS["crystal_depth"], FS["radius"], FS["dets_per_ring"], FS["num_rings"],
                    FS["num_doi"], FS["max_ring_diff"], FS["min_ang_diff"], FS["dets_per_block"])


def test_lut_closed_form():
    sc = O.ScannerDef(96.0, 4.0, 4.0, 20.0, 300.0, 256, 24)
    pos, ori = O.lut(sc)
    r = np.hypot(pos[:, 0], pos[:, 1])
    np.testing.assert_allclose(r, 310.0, rtol=2e-7)
    ring = np.arange(sc.num_dets) // 256
    np.testing.assert_allclose(pos[:, 2], -48 + (ring + 0.5) * 4.0, atol=1e-5)
    ang = 2 * np.pi * (np.arange(sc.num_dets) % 256) / 256
    np.testing.assert_allclose(ori[:, 0], np.cos(ang), atol=1e-6)
    np.testing.assert_allclose(ori[:, 1], np.sin(ang), atol=1e-6)
    # blocks of several crystals: crystals of a block lie on the block face
    pos, ori = O.lut(FAKE)
    dpb = FAKE.dets_per_block
    for b in range(FAKE.dets_per_ring // dpb):
        blk = pos[b * dpb:(b + 1) * dpb]
        n = ori[b * dpb]
        np.testing.assert_allclose(blk[:, :2] @ n[:2], FAKE.radius + FAKE.crystal_depth / 2, rtol=1e-6)


def test_histo_dims_config1():
    sc = O.ScannerDef(96.0, 4.0, 4.0, 20.0, 300.0, 256, 24, max_ring_diff=23)
    assert O.histo_dims(sc) == (576, 256, 128)


def test_histo_roundtrip_fake_scanner():
    """test_Histogram3D.cpp:140-300: bin -> pair -> bin round trip over every
    bin, pairs within the LUT, no duplicate pairs."""
    zb, ph, nr = O.histo_dims(FAKE)
    bins = np.arange(zb * ph * nr, dtype=np.int64)
    d1, d2 = O.histo_pairs(FAKE, bins)
    nd = FAKE.num_dets
    assert (d1 < nd).all() and (d2 < nd).all()
    back = O.histo_bins_from_pairs(FAKE, d1, d2)
    np.testing.assert_array_equal(back, bins)
    key = np.minimum(d1, d2).astype(np.int64) * nd + np.maximum(d1, d2)
    assert np.unique(key).size == key.size


def test_histo_allowed_lors_fake_scanner():
    """test_Histogram3D.cpp:331-375: every allowed pair (ring difference <=
    MRD, in-ring angle difference >= minAngDiff) maps to a bin whose pair is
    itself."""
    dpr, nr, nd = FAKE.dets_per_ring, FAKE.num_rings, FAKE.num_dets
    rng = np.random.default_rng(0)
    d1 = rng.integers(0, nd, 20000)
    d2 = rng.integers(0, nd, 20000)
    r1, r2 = (d1 // dpr) % nr, (d2 // dpr) % nr
    c1, c2 = d1 % dpr, d2 % dpr
    ad = np.minimum((c1 - c2) % dpr, (c2 - c1) % dpr)
    ok = (np.abs(r1 - r2) <= FAKE.max_ring_diff) & (ad >= FAKE.min_ang_diff) & (d1 != d2)
    b = O.histo_bins_from_pairs(FAKE, d1[ok], d2[ok])
    assert (b >= 0).all()
    e1, e2 = O.histo_pairs(FAKE, b)
    same = ((e1 == d1[ok]) & (e2 == d2[ok])) | ((e1 == d2[ok]) & (e2 == d1[ok]))
    assert same.all()


def test_histo_subsets_partition():
    """BinIteratorRangeHistogram3D: the subsets partition the bins, each
    subset holds whole phi planes phi = S*k + s (test_Histogram3D.cpp:
    300-318)."""
    zb, ph, nr = O.histo_dims(FAKE)
    for S in (1, 5, 7, 48):
        allb = np.concatenate([O.histo_subset_bins(FAKE, S, s) for s in range(S)])
        assert allb.size == zb * ph * nr
        assert np.unique(allb).size == allb.size
        for s in range(S):
            b = O.histo_subset_bins(FAKE, S, s)
            phi = (b % (ph * nr)) // nr
            assert ((phi % S) == s).all()
    # one subset per angle: size = numR * numZBin
    assert O.histo_subset_bins(FAKE, ph, 3).size == nr * zb


def test_listmode_subset_sizes():
    """BinIteratorChronologicalInterleaved (test_BinIterator.cpp:82-130):
    size (N - 1 - s) / S + 1, start s, stride S."""
    for N in (1, 13, 100, 101):
        for S in (1, 3, 10):
            tot = 0
            for s in range(min(S, N)):
                idx = O.listmode_subset(N, S, s)
                assert idx.size == (N - 1 - s) // S + 1
                assert (np.diff(idx) == S).all() and idx[0] == s
                tot += idx.size
            assert tot == N


def test_em_update_and_threshold_closed_forms():
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 2, 1000).astype(np.float32)
    u = rng.uniform(0, 2, 1000).astype(np.float32)
    s = rng.uniform(-1, 2, 1000).astype(np.float32)
    thr = np.float32(0.2)
    ref = np.where(s > thr, x * (u / s), x).astype(np.float32)
    got = O.em_update(x.copy(), u, s, thr)
    np.testing.assert_array_equal(got, ref)
    t = O.apply_threshold(x.copy(), s, 0.5, 0.0, 0.1, 2.0, 0.3)
    np.testing.assert_array_equal(t, np.where(s <= 0.5, x * np.float32(0.0) + np.float32(0.1),
                                              x * np.float32(2.0) + np.float32(0.3)).astype(np.float32))


def _small_problem(tof=False, rays=1, seed=2, n=3000):
    sc = O.ScannerDef(64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    pos, ori = O.lut(sc)
    im = O.ImageDef(24, 24, 12, 6.0, 6.0, 5.0)
    rng = np.random.default_rng(seed)
    d1 = rng.integers(0, sc.num_dets, n)
    d2 = rng.integers(0, sc.num_dets, n)
    L = O.lors_from_pairs(pos, ori, d1, d2, tof=rng.uniform(-300, 300, n) if tof else None)
    p = O.proj_for_scanner(sc, num_rays=rays, tof_width_ps=300.0 if tof else None, tof_n_std=5)
    return sc, im, L, p, rng


@pytest.mark.parametrize("tof,rays", [(False, 1), (True, 4)])
def test_mlem_count_preservation(tof, rays):
    """After one MLEM update with S = A^T 1 over the same LORs and no
    corrections: sum_j x_j S_j = sum_i m_i over the sensitive voxels."""
    sc, im, L, p, rng = _small_problem(tof, rays)
    x = rng.uniform(0.5, 1.5, im.shape)
    m = rng.poisson(3.0, L.n).astype(np.float64)
    S = O.back(im, p, L, np.ones(L.n), dtype=np.float64)[0]
    ax = O.forward(im, x, p, L, dtype=np.float64)
    keep = ax > 1e-35
    u = np.zeros(im.shape)
    O.em_subset(im, x, u, p, L, meas=m, dtype=np.float64)
    xn = O.em_update(x.copy(), u, S, np.float64(O.EPS_FLT))
    sens = S > O.EPS_FLT
    # exact up to the mass landing in voxels with S <= EPS_FLT (not updated):
    # with TOF tails that is ~1e-9 of the total
    assert (xn[0][sens] * S[sens]).sum() == pytest.approx(m[keep].sum(), rel=1e-8)
    lost = (u[0][~sens] * x[0][~sens]).sum()
    assert (xn[0][sens] * S[sens]).sum() + lost == pytest.approx(m[keep].sum(), rel=1e-12)
    assert (xn >= 0).all()


def test_mlem_fixed_point():
    """Consistent data m = A x*: x* is a fixed point of the EM update."""
    sc, im, L, p, rng = _small_problem(True, 4)
    xs = rng.uniform(0.5, 1.5, im.shape)
    m = O.forward(im, xs, p, L, dtype=np.float64)
    S = O.back(im, p, L, np.ones(L.n), dtype=np.float64)[0]
    u = np.zeros(im.shape)
    O.em_subset(im, xs, u, p, L, meas=m, dtype=np.float64)
    xn = O.em_update(xs.copy(), u, S, np.float64(O.EPS_FLT))
    sens = S > O.EPS_FLT
    np.testing.assert_allclose(xn[0][sens], xs[0][sens], rtol=1e-10)


def test_em_subset_one_lor_by_hand():
    """One LOR, one correction set: u = A^T( m / (A x * a * g + r) * a * g )
    (OSEM_CPU.cpp:329-374) computed by hand from forward / back."""
    sc, im, L, p, rng = _small_problem(False, 1, n=40)
    x = rng.uniform(0.5, 1.5, im.shape)
    ax = O.forward(im, x, p, L, dtype=np.float64)
    i = int(np.argmax(ax))
    L1 = L.take(np.array([i]))
    a, g, r, m = 0.7, 1.3, 0.05, 4.0
    u = np.zeros(im.shape)
    O.em_subset(im, x, u, p, L1, meas=[m], att=[a], rnd=[r], global_scale=g, dtype=np.float64)
    v = m / (ax[i] * a * g + r) * a * g
    ref = O.back(im, p, L1, np.array([v]), dtype=np.float64)
    np.testing.assert_allclose(u, ref, rtol=1e-12)


def test_osem_flow_invariants():
    """Restated OSEM driver on a small histogram: positivity, mask outside
    the sensitive region, and likelihood increase over iterations."""
    sc = O.ScannerDef(48.0, 4.0, 4.0, 10.0, 110.0, 64, 12, max_ring_diff=7)
    im = O.ImageDef(16, 16, 12, 6.0, 6.0, 4.0)
    pos, ori = O.lut(sc)
    p = O.proj_for_scanner(sc)
    zb, ph, nr = O.histo_dims(sc)
    bins = np.arange(zb * ph * nr)
    d1, d2 = O.histo_pairs(sc, bins)
    L = O.lors_from_pairs(pos, ori, d1, d2)
    rng = np.random.default_rng(3)
    xt = np.zeros(im.shape, np.float32)
    xt[0, 3:9, 5:11, 5:11] = 2.0
    y = rng.poisson(O.forward(im, xt, p, L) * 5).astype(np.float32)
    S = 4
    sens = O.sensitivity_images(sc, im, p, S, False, pos_ori=(pos, ori))
    subsets = [O.histo_subset_bins(sc, S, s) for s in range(S)]
    lls = []
    x = None
    for it in range(3):
        x = O.reconstruct(sc, im, p, L, sens, 1, S, False, meas=y, subset_indices=subsets,
                          initial=x)
        ax = O.forward(im, x, p, L).astype(np.float64)
        ok = ax > 0
        lls.append(float((y[ok] * np.log(ax[ok]) - ax[ok]).sum()))
    assert (x >= 0).all()
    assert lls[1] > lls[0] and lls[2] > lls[1]
