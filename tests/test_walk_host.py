"""CPU checks of the device Siddon walk (yrt-pet_amd/csrc/yrtmi_rays.hpp),
compiled for the host with tests/hostwalk/ (a qualifier shim; test only).

1. The library's own walk (setup + traversal, both setup precisions, TOF)
   against the oracle's restatement of ProjectorSiddon::project_helper:
   same segment count per ray, projections <= 1e-6 relative L2 (FP32 setup
   vs the FP32 oracle) / <= 1e-6 (FP64 setup vs the FP64 oracle).
2. The tile decomposition used by tile-planned backprojection: every ray's
   walk cut at tile boundaries and each piece re-entered from its tile-entry
   record (sidPlaneShift + sidRestart) must add up to the full walk:
   bit-exact per voxel on voxel sizes that are short binary fractions, and
   within 3e-4 of the ray's largest voxel weight otherwise: plane positions
   recomputed directly vs the walk's accumulated ones differ by an ulp, which
   moves a segment end by ~ ulp(alpha) * ray length = 6e-8 * 400 mm = 2.4e-5
   mm per end against voxel weights of ~1 mm (measured worst 1.03e-4 with
   8x4x2 tiles, the most restarts per ray).
   The event sets include the rays that exposed the two restart cases: a
   first segment an ulp long at a plane corner (the walk's first voxel then
   sits one index over for the whole ray) and a plane recomputed an ulp
   before the tile entry.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import yrtmi
from helpers import oracle_image, oracle_proj, oracle_scanner, rel_l2
from yrtmi import synth

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def hw(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hostwalk") / "libhostwalk.so")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC",
           "-I", os.path.join(HERE, "hostwalk"),
           "-I", os.path.join(REPO, "yrt-pet_amd", "csrc"),
           os.path.join(HERE, "hostwalk", "walk_host.cpp"), "-o", out]
    subprocess.run(cmd, check=True, capture_output=True)
    lib = C.CDLL(out)
    P = C.c_void_p
    lib.hw_forward.restype = C.c_float
    lib.hw_forward.argtypes = [P, P, P, P, C.c_int, C.c_int, P, C.c_float, C.POINTER(C.c_int)]
    lib.hw_tiles.restype = C.c_double
    lib.hw_tiles.argtypes = [P, P, P, C.c_int, C.c_int, P, C.c_float, C.c_int, C.c_int,
                             C.c_int, P]
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def imp_of(im):
    lx, ly, lz = im.lengths()
    return np.array([im.nx, im.ny, im.nz, im.vx, im.vy, im.vz, lx, ly, lz, im.fovRadius],
                    np.float32)


def rays(sc, n, seed, tof, R):
    """Nominal and multi-ray lines of n random events (oracle geometry)."""
    ev = synth.random_events(sc, n, np.random.default_rng(seed), tof_fwhm_ps=tof)
    pos, ori = O.lut(oracle_scanner(sc))
    L = O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev.get("tof"))
    p1, p2, tv = [], [], []
    for i in range(L.n):
        t = float(L.tof[i]) if L.tof is not None else 0.0
        p1.append(L.p1[i]); p2.append(L.p2[i]); tv.append(t)
        if R > 1:
            ml = O.multiray_lines(L.p1[i], L.p2[i], L.n1[i], L.n2[i], sc.crystalSize_z,
                                  sc.crystalSize_trans, R - 1)
            for k in range(R - 1):
                p1.append(ml[k, :3]); p2.append(ml[k, 3:]); tv.append(t)
    f = lambda a: np.ascontiguousarray(np.array(a, np.float32))
    return f(p1), f(p2), np.array(tv, np.float32), ev, L


SCANNER = dict(t=("t", 96.0, 4.0, 4.0, 10.0, 180.0, 128, 24),
               small=("small", 64.0, 4.0, 4.0, 10.0, 150.0, 96, 16))


@pytest.mark.parametrize("tof,s64", [(None, 0), (250.0, 0), (250.0, 1)])
def test_host_walk_matches_oracle(hw, tof, s64):
    sc = yrtmi.Scanner(*SCANNER["small"])
    im = yrtmi.ImageParams(40, 36, 16, 5.0, 5.5, 4.0)
    p1, p2, tv, ev, L = rays(sc, 1500, 11, tof, 1)
    x = np.random.default_rng(5).uniform(0, 2, im.shape).astype(np.float32)
    pp = yrtmi.ProjectorParams(sc)
    if tof:
        pp.addTOF(tof, 5)
    dt = np.float64 if s64 else np.float32
    ref = O.forward(oracle_image(im), x.astype(dt), oracle_proj(pp), L, dtype=dt)
    cnt_ref = O.count_segments(oracle_image(im), oracle_proj(pp), L)
    sig, tr, nm = O.tof_helper(tof or 250.0, 5)
    tofc = np.array([sig, tr, nm], np.float32)
    imp = imp_of(im)
    got = np.zeros(L.n, np.float32)
    cnt = np.zeros(L.n, np.int64)
    c = C.c_int()
    for i in range(L.n):
        got[i] = hw.hw_forward(_p(imp), _p(x), _p(p1[i]), _p(p2[i]), 1 if tof else 0, s64,
                               _p(tofc), float(tv[i]), C.byref(c))
        cnt[i] = c.value
    r = rel_l2(got, ref)
    print(f"host walk tof={tof} s64={s64}: rel L2 {r:.2e}, segments {cnt.sum()} vs {cnt_ref.sum()}")
    assert r <= 1e-6
    if not s64:
        assert (cnt == cnt_ref).all()


TILE_CASES = [
    # (name, scanner, image, seed, n, tof, s64, rays, exact)
    ("dyadic_r1", "t", (70, 66, 40, 4.0, 4.5, 2.25), 103, 8000, None, 0, 1, True),
    ("dyadic_r4_tof", "t", (70, 66, 40, 4.0, 4.5, 2.25), 101, 2000, 250.0, 0, 4, True),
    ("dyadic_r4_tof_fp64", "t", (70, 66, 40, 4.0, 4.5, 2.25), 101, 2000, 250.0, 1, 4, True),
    ("nondyadic_r4_tof_fp64", "t", (61, 57, 35, 5.1, 5.3, 2.7), 101, 6000, 300.0, 1, 4, False),
    ("nondyadic_r3", "t", (61, 57, 35, 5.1, 5.3, 2.7), 101, 2000, None, 0, 3, False),
]


@pytest.mark.parametrize("case", TILE_CASES, ids=[c[0] for c in TILE_CASES])
def test_tile_decomposition(hw, case):
    name, scn, imd, seed, n, tof, s64, R, exact = case
    sc = yrtmi.Scanner(*SCANNER[scn])
    im = yrtmi.ImageParams(*imd)
    p1, p2, tv, _, _ = rays(sc, n, seed, tof, R)
    sig, tr, nm = O.tof_helper(tof or 250.0, 5)
    tofc = np.array([sig, tr, nm], np.float32)
    imp = imp_of(im)
    out = np.zeros(4, np.int32)
    worst, tiles, segs, stored = 0.0, 0, 0, 0
    for (tx, ty, tz) in ((32, 32, 16), (8, 4, 2)):
        for i in range(len(p1)):
            d = hw.hw_tiles(_p(imp), _p(p1[i]), _p(p2[i]), 1 if tof else 0, s64, _p(tofc),
                            float(tv[i]), tx, ty, tz, _p(out))
            worst = max(worst, d)
            tiles += int(out[0])
            segs += int(out[1])
            stored += int(out[3])
    print(f"{name}: {len(p1)} rays x 2 tile shapes, {tiles} tile entries, {segs} segments, "
          f"worst per-voxel deviation {worst:.2e}, stored-set-up mismatches {stored}")
    assert segs > 0 and tiles > len(p1)
    # re-entering a tile from the plan's stored FP32 set-up walks exactly as
    # from a fresh set-up
    assert stored == 0
    if exact:
        assert worst == 0.0
    else:
        assert worst <= 3e-4


def bricked(x):
    """The forward's bricked copy of an x-major frame (4 x 4 x 2 voxel bricks,
    x-major inside and across bricks, padded with zeros)."""
    nz, ny, nx = x.shape
    bx, by, bz = (nx + 3) // 4, (ny + 3) // 4, (nz + 1) // 2
    p = np.zeros((bz * 2, by * 4, bx * 4), np.float32)
    p[:nz, :ny, :nx] = x
    # (z, y, x) -> (bz, 2, by, 4, bx, 4) -> (bz, by, bx, 2, 4, 4)
    return np.ascontiguousarray(p.reshape(bz, 2, by, 4, bx, 4).transpose(0, 2, 4, 1, 3, 5)).reshape(-1)


@pytest.mark.parametrize("tof", [None, 300.0])
def test_bricked_forward_matches_row_major(hw, tof):
    """The forward over the bricked frame visits the same voxels in the same
    order as over the x-major frame: every segment's bricked index is its
    voxel's, the index fetched a segment ahead is the next voxel's, and the
    line integrals are bit-identical (odd sizes: partial bricks)."""
    P = C.c_void_p
    hw.hw_brick.restype = C.c_int
    hw.hw_brick.argtypes = [P, P, P, P, P, C.c_int, P, C.c_float, P]
    sc = yrtmi.Scanner(*SCANNER["t"])
    im = yrtmi.ImageParams(61, 57, 35, 2.1, 2.3, 2.7)
    p1, p2, tv, _, _ = rays(sc, 3000, 11, tof, 4)
    sig, tr, nm = O.tof_helper(tof or 250.0, 5)
    tofc = np.array([sig, tr, nm], np.float32)
    imp = imp_of(im)
    x = np.random.default_rng(3).uniform(0.5, 2.0, (im.nz, im.ny, im.nx)).astype(np.float32)
    xb = bricked(x)
    out = np.zeros(2, np.float32)
    bad, n = 0, 0
    for i in range(len(p1)):
        bad += hw.hw_brick(_p(imp), _p(x), _p(xb), _p(p1[i]), _p(p2[i]), 1 if tof else 0,
                           _p(tofc), float(tv[i]), _p(out))
        assert out[0] == out[1], (i, out)
        n += out[1] != 0
    print(f"bricked forward tof={tof}: {len(p1)} rays ({n} nonzero), index mismatches {bad}")
    assert bad == 0 and n > 1000


# ---- distance-driven sweep (ddSetup + ddVisit) --------------------------
DD_CASES = [
    # (name, tof, psf): unequal crystal sizes, anisotropic voxels
    ("dd", None, False),
    ("dd_tof", 300.0, False),
    ("dd_psf", None, True),
    ("dd_psf_tof", 300.0, True),
]


@pytest.mark.parametrize("case", DD_CASES, ids=[c[0] for c in DD_CASES])
def test_host_dd_sweep_matches_oracle(hw, case):
    """The library's DD sweep (per-LOR reciprocals, z weights once per slice,
    the PSF integral from the prefix table) against the oracle's restatement
    of ProjectorDD::dd_project_ref: the same footprint voxels, forward
    projections and backprojections within FP32 rounding (<= 1e-6 relative
    L2; the PSF table sums whole kernel segments in FP64 where the reference
    adds them one by one in FP32, <= 2e-6), and a sweep cut into tiles adds
    up to the whole sweep exactly."""
    name, tof, psf = case
    lib = hw
    P = C.c_void_p
    lib.hw_dd.restype = C.c_float
    lib.hw_dd.argtypes = [P, P, P, P, P, P, C.c_int, P, C.c_float, C.c_float, C.c_float, P,
                          C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, P, P, P,
                          C.POINTER(C.c_int)]
    sc = yrtmi.Scanner("dd", 64.0, 4.0, 3.2, 10.0, 150.0, 96, 16)
    im = yrtmi.ImageParams(37, 33, 19, 4.2, 3.7, 3.3)
    rng = np.random.default_rng(17)
    ev = synth.random_events(sc, 700, rng, tof_fwhm_ps=tof)
    osc = oracle_scanner(sc)
    pos, ori = O.lut(osc)
    L = O.lors_from_pairs(pos, ori, ev["d1"], ev["d2"], tof=ev.get("tof"))
    k = s_step = ks = None
    taps = None
    if psf:
        # two asymmetric rows (the flip matters for backprojection)
        K = 21
        g = np.exp(-0.5 * ((np.arange(K) - 8.0) / 3.0) ** 2)
        taps = np.stack([g / g.sum() / 0.25, (g[::-1] ** 1.5) / (g ** 1.5).sum() / 0.25])
        taps = taps.astype(np.float32)
        k, s_step, ks = taps, 60.0, 0.25
    po = O.proj_for_scanner(osc, type="dd", tof_width_ps=tof, tof_n_std=5 if tof else 0,
                            psf_kernels=k, psf_s_step=s_step or 0.0, psf_k_spacing=ks or 0.0)
    x = rng.uniform(0.2, 2.0, im.shape).astype(np.float32)
    oi = oracle_image(im)
    ref = O.forward(oi, x, po, L)
    cnt_ref = O.count_segments(oi, po, L)
    vals = rng.uniform(0.5, 1.5, L.n).astype(np.float32)
    bref = O.back(oi, po, L, vals)
    sig, tr, nm = O.tof_helper(tof or 250.0, 5)
    tofc = np.array([sig, tr, nm], np.float32)
    imp = imp_of(im)
    xs = np.ascontiguousarray(x.reshape(-1))
    got = np.zeros(L.n, np.float32)
    cnt = np.zeros(L.n, np.int64)
    back = np.zeros(im.nvox, np.float32)
    c = C.c_int()
    tap_p = _p(np.ascontiguousarray(taps)) if taps is not None else None
    worst_tile = 0.0
    tiles = np.array([8, 4, 2], np.int32)
    for i in range(L.n):
        a = np.ascontiguousarray(L.p1[i], np.float32)
        b = np.ascontiguousarray(L.p2[i], np.float32)
        n1 = np.ascontiguousarray(L.n1[i], np.float32)
        n2 = np.ascontiguousarray(L.n2[i], np.float32)
        one = np.zeros(im.nvox, np.float32)
        til = np.zeros(im.nvox, np.float32)
        tv = float(L.tof[i]) if L.tof is not None else 0.0
        got[i] = lib.hw_dd(_p(imp), _p(xs), _p(a), _p(b), _p(n1), _p(n2), 1 if tof else 0,
                           _p(tofc), tv, sc.crystalSize_z, sc.crystalSize_trans, tap_p,
                           2 if psf else 0, 21 if psf else 0, s_step or 0.0, ks or 0.0,
                           float(vals[i]), _p(one), _p(tiles) if i % 7 == 0 else None,
                           _p(til) if i % 7 == 0 else None, C.byref(c))
        cnt[i] = c.value
        back += one
        if i % 7 == 0:
            worst_tile = max(worst_tile, float(np.abs(til - one).max()))
    tol = 2e-6 if psf else 1e-6
    rf, rb = rel_l2(got, ref), rel_l2(back.reshape(im.shape), bref)
    print(f"{name}: forward rel L2 {rf:.2e}, back {rb:.2e}, voxels {cnt.sum()} vs "
          f"{cnt_ref.sum()}, tiles worst {worst_tile:.1e}")
    assert (cnt == cnt_ref).all()
    assert rf <= tol and rb <= tol
    assert worst_tile == 0.0
