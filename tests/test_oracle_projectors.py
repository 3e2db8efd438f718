"""Pins of the oracle's Siddon and distance-driven projectors.

Siddon: the closed forms, regression coordinates and adjoint identities of
the reference's test_Siddon.cpp, plus an FP64 brute-force voxel-intersection
check (the reference uses a 2 % bar; the FP64 oracle is held to 1e-9 and the
FP32 restatement to 1e-4).  DD: getOverlap known answers (test_DD.cpp),
uniform-image slab identity, adjoint, FP32 vs FP64.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "reference_known_answers.json")))


def lor(p1, p2, n1=None, n2=None, tof=None):
    a = lambda v: np.asarray([v], np.float32)
    return O.Lors(a(p1), a(p2), None if n1 is None else a(n1), None if n2 is None else a(n2),
                  None if tof is None else np.asarray([tof], np.float32))


def fwd1(im, img, L, dtype=np.float32, proj=None):
    return float(O.forward(im, img, proj or O.ProjDef(), L, dtype=dtype)[0])


SIMPLE = KA["siddon_simple_image"]
IM_S = O.ImageDef.from_lengths(*SIMPLE["image_n"], *SIMPLE["image_length"])
ONES = np.ones(IM_S.shape, np.float32)
R = float(IM_S.fov_radius)
SX, SY, SZ = SIMPLE["image_length"]


@pytest.mark.parametrize("i", range(10))
def test_siddon_isocenter_ray(i):
    b = 2 * math.pi * i / 9
    L = lor([-SX * math.cos(b), -SX * math.sin(b), 0], [SX * math.cos(b), SX * math.sin(b), 0])
    assert fwd1(IM_S, ONES, L) == pytest.approx(2 * R, rel=1e-5)
    assert fwd1(IM_S, ONES.astype(np.float64), L, np.float64) == pytest.approx(2 * R, rel=1e-12)


def test_siddon_within_fov_and_chords():
    rng = np.random.default_rng(1)
    for _ in range(20):
        b1, b2 = rng.uniform(0, 2 * math.pi, 2)
        r1, r2 = rng.uniform(0, 0.8 * R, 2)
        p1 = [r1 * math.cos(b1), r1 * math.sin(b1), 0]
        p2 = [r2 * math.cos(b2), r2 * math.sin(b2), 0]
        ref = float(np.linalg.norm(np.float32(p1) - np.float32(p2)))
        assert fwd1(IM_S, ONES, lor(p1, p2)) == pytest.approx(ref, rel=1e-5)
    for i in range(10):
        y0 = SY * i / 9 - SY / 2
        ref = 2 * math.sqrt(max(0.0, R * R - y0 * y0))
        assert fwd1(IM_S, ONES, lor([-SX, y0, 0], [SX, y0, 0])) == pytest.approx(ref, rel=1e-5, abs=1e-6)


def test_siddon_outside_rays():
    assert fwd1(IM_S, ONES, lor([SX, 0, 0], [2 * SX, 0, 0])) == 0.0
    assert fwd1(IM_S, ONES, lor([2 * SX, 0, 0], [2 * SX, SY, 0])) == 0.0
    assert fwd1(IM_S, ONES, lor([-SX, 0, 1.0001 * SZ / 2], [SX, 0, 1.0001 * SZ / 2])) == 0.0


def test_siddon_z_rays():
    rng = np.random.default_rng(2)
    for _ in range(10):
        z1, z2 = rng.uniform(-SZ / 2, SZ / 2, 2)
        ref = math.sqrt(4 * R * R + (z2 - z1) ** 2)
        assert fwd1(IM_S, ONES, lor([0, -R, z1], [0, R, z2])) == pytest.approx(ref, rel=1e-5)
    cases = [([0, 0, SZ], [0, 0, -SZ], SZ), ([0, -SY, 0], [0, SY, 0], 2 * R),
             ([-SX, 0, 0], [SX, 0, 0], 2 * R), ([-SX, -SY, 0], [SX, SY, 0], 2 * R)]
    for p1, p2, ref in cases:
        assert fwd1(IM_S, ONES, lor(p1, p2)) == pytest.approx(ref, rel=1e-5)


def test_siddon_regressions():
    k = KA["siddon_bug_offset"]
    im = O.ImageDef.from_lengths(*k["image_n"], *k["image_length"])
    v = 123.25
    img = np.full(im.shape, v, np.float32)
    assert fwd1(im, img, lor(k["p1"], k["p2"])) == pytest.approx(v * k["expected_over_value"], rel=1e-5)
    k = KA["siddon_bug_zext"]
    im = O.ImageDef.from_lengths(*k["image_n"], *k["image_length"])
    img = np.full(im.shape, v, np.float32)
    f32 = fwd1(im, img, lor(k["p1"], k["p2"]))
    f64 = fwd1(im, img.astype(np.float64), lor(k["p1"], k["p2"]), np.float64)
    assert f32 > 0 and f32 == pytest.approx(f64, rel=1e-5)
    k = KA["siddon_multi_intersection"]
    im = O.ImageDef.from_lengths(*k["image_n"], *k["image_length"])
    img = np.random.default_rng(3).uniform(-5, 5, im.shape).astype(np.float32)
    for p1, p2 in k["lors"]:
        assert fwd1(im, img, lor(p1, p2)) == pytest.approx(
            fwd1(im, img.astype(np.float64), lor(p1, p2), np.float64), rel=1e-5, abs=1e-5)


def brute_force(im, img, p1, p2):
    """Exact line integral by intersecting the segment with every voxel and
    the FOV cylinder, in FP64 (test_Siddon.cpp:389-500 restated)."""
    p1 = np.asarray(p1, np.float64)
    p2 = np.asarray(p2, np.float64)
    lx, ly, lz = (float(v) for v in im.lengths)
    R = float(im.fov_radius)
    d = p2 - p1
    A = d[0] ** 2 + d[1] ** 2
    B = 2 * (d[0] * p1[0] + d[1] * p1[1])
    Cc = p1[0] ** 2 + p1[1] ** 2 - R * R
    if A != 0:
        D = B * B - 4 * A * Cc
        if D <= 0:
            return 0.0
        t1, t2 = (-B - math.sqrt(D)) / (2 * A), (-B + math.sqrt(D)) / (2 * A)
    else:
        t1, t2 = 0.0, 1.0
    t1, t2 = max(0.0, t1), min(1.0, t2)
    dn = float(np.linalg.norm(d))
    if dn == 0 or t1 >= t2:
        return 0.0
    # voxel planes at -length/2 + k * voxel_size, the voxel size being the
    # stored FP32 value (ProjectorSiddon.cpp:532-544)
    # and the outer faces at +-length/2 (the box clip, 546-553).  The planes
    # are counted from the face the ray enters (x_cur = inv > 0 ? x0 : x1,
    # 564-566); as n * v != length in FP32 they sit ~1 ulp apart for the two
    # directions, which this restates.
    edges = []
    for ax, (l, n, v) in enumerate(((lx, im.nx, im.vx), (ly, im.ny, im.vy), (lz, im.nz, im.vz))):
        v = float(np.float32(v))
        # the exit face is whichever comes first of the last counted plane and
        # the box face (the traversal leaves the grid at either)
        if d[ax] >= 0:
            e = np.concatenate([-l / 2 + np.arange(n) * v, [min(l / 2, -l / 2 + n * v)]])
        else:
            e = np.concatenate([[max(-l / 2, l / 2 - n * v)], l / 2 - np.arange(n - 1, -1, -1) * v])
        edges.append(e)
    lo, hi = [], []
    for ax in range(3):
        e = edges[ax]
        if d[ax] != 0:
            a0 = (e[:-1] - p1[ax]) / d[ax]
            a1 = (e[1:] - p1[ax]) / d[ax]
            lo.append(np.minimum(a0, a1))
            hi.append(np.maximum(a0, a1))
        else:
            inside = (e[:-1] <= p1[ax]) & (p1[ax] < e[1:])
            lo.append(np.where(inside, -np.inf, np.inf))
            hi.append(np.where(inside, np.inf, -np.inf))
    amin = np.maximum(np.maximum(lo[2][:, None, None], lo[1][None, :, None]), lo[0][None, None, :])
    amax = np.minimum(np.minimum(hi[2][:, None, None], hi[1][None, :, None]), hi[0][None, None, :])
    amin = np.maximum(amin, t1)
    amax = np.minimum(amax, t2)
    w = np.clip(amax - amin, 0, None) * dn
    return float((w * img[0]).sum())


def test_siddon_brute_force():
    rng = np.random.default_rng(4)
    im = O.ImageDef.from_lengths(13, 11, 7, 9.1, 7.3, 5.2)
    img = rng.uniform(-5, 5, im.shape)
    lx, ly, lz = (float(v) for v in im.lengths)
    n = 0
    for _ in range(150):
        p1 = rng.uniform(-1, 1, 3) * [lx, ly, lz]
        p2 = rng.uniform(-1, 1, 3) * [lx, ly, lz]
        p1, p2 = np.float32(p1), np.float32(p2)
        ref = brute_force(im, img, p1, p2)
        f64 = fwd1(im, img, lor(p1, p2), np.float64)
        f32 = fwd1(im, img.astype(np.float32), lor(p1, p2))
        assert f64 == pytest.approx(ref, rel=1e-9, abs=1e-9)
        assert f32 == pytest.approx(ref, rel=1e-4, abs=1e-4)
        n += ref != 0
    assert n > 50


def test_siddon_adjoint_multiray_tof():
    rng = np.random.default_rng(5)
    sc = O.ScannerDef(64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    pos, ori = O.lut(sc)
    im = O.ImageDef(40, 36, 16, 5.0, 5.5, 4.0, 2.0, -3.0, 1.0)
    n = 2000
    d1 = rng.integers(0, sc.num_dets, n)
    d2 = rng.integers(0, sc.num_dets, n)
    L = O.lors_from_pairs(pos, ori, d1, d2, tof=rng.uniform(-500, 500, n))
    for proj in (O.proj_for_scanner(sc, num_rays=4, tof_width_ps=250.0, tof_n_std=5),
                 O.proj_for_scanner(sc, type="dd", tof_width_ps=300.0, tof_n_std=4),
                 O.proj_for_scanner(sc, type="dd")):
        x = rng.uniform(0, 1, im.shape)
        y = rng.uniform(0, 1, n)
        ax = O.forward(im, x, proj, L, dtype=np.float64)
        aty = O.back(im, proj, L, y, dtype=np.float64)
        assert abs(ax @ y - (x * aty).sum()) <= 1e-12 * abs(ax @ y)


def test_dd_get_overlap_known_answers():
    for p0, p1, d0, d1, ref in KA["dd_get_overlap"]["cases"]:
        assert O.dd_overlap(p0, p1, d0, d1) == pytest.approx(ref, abs=1e-6)


def test_dd_uniform_image_slab_identity():
    """For a LOR whose footprint stays inside the image, DD forward of a
    uniform image sums dxy/|cos(theta)| over the swept slices
    (ProjectorDD.cpp:437-445,533): the path length between the first and the
    last slice centres plus one slice."""
    sc = O.ScannerDef(64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    pos, ori = O.lut(sc)
    im = O.ImageDef(64, 64, 32, 4.0, 4.0, 4.0)
    ones = np.ones(im.shape, np.float64)
    # two detectors facing each other in the same ring, near x = 0
    d1, d2 = 16 * 96 // 2 + 24, 16 * 96 // 2 + 72  # ring 8, angles 90 / 270 deg
    L = O.lors_from_pairs(pos, ori, [d1], [d2])
    p = O.proj_for_scanner(sc, type="dd")
    v = O.forward(im, ones, p, L, dtype=np.float64)[0]
    dy = abs(float(L.p1[0, 1] - L.p2[0, 1]))
    cos_t = dy / float(np.linalg.norm(L.p1[0] - L.p2[0]))
    assert v == pytest.approx(im.ny * im.vy / cos_t, rel=1e-6)


def test_dd_fp32_vs_fp64_and_psf():
    rng = np.random.default_rng(6)
    sc = O.ScannerDef(64.0, 4.0, 4.0, 10.0, 150.0, 96, 16)
    pos, ori = O.lut(sc)
    im = O.ImageDef(40, 36, 16, 5.0, 5.5, 4.0)
    n = 1000
    L = O.lors_from_pairs(pos, ori, rng.integers(0, sc.num_dets, n), rng.integers(0, sc.num_dets, n))
    x = rng.uniform(0, 2, im.shape)
    p = O.proj_for_scanner(sc, type="dd")
    f32 = O.forward(im, x.astype(np.float32), p, L)
    f64 = O.forward(im, x, p, L, dtype=np.float64)
    assert np.linalg.norm(f32 - f64) / np.linalg.norm(f64) < 1e-2
    # wide flat PSF integrates over its support: the PSF weight of an
    # interval fully inside a flat unit-density kernel equals its length
    k = np.full((1, 81), 1.0 / 8.0, np.float32)  # density 1/8 over 20 mm
    assert O.psf_integral(k[0], 0.25, -2.0, 2.0) == pytest.approx(0.5, rel=1e-6)
