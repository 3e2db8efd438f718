"""Closed-form pins of the oracle's distance-driven (DD) projector.

The weights the DD projector gives a voxel (ProjectorDD.cpp:260-578) are,
for a LOR whose detectors face each other along a coordinate axis, a product
of closed forms:

* transaxial: the overlap of the voxel's minor-axis extent with the detector
  footprint, the segment between the two crystals' transaxial edges
  (crystal_size_trans wide), divided by the footprint width (474-492);
* axial: the same for z with the crystals' axial edges (crystal_size_z wide),
  the footprint centred on the LOR's z at that slice (493-531);
* the slice's path length: slice thickness / |cos theta| (437-445, 533);
* TOF: the Gaussian (TimeOfFlight.hpp:74-82) at the slab centre's distance
  from the first detector, the slab being the voxel slice (500-515);
* projection PSF: the integral of the kernel's linear interpolant (with one
  zero sample of padding each side, ProjectionPsfManager.cpp:52-56) over
  [voxel low - footprint high, voxel high - footprint low] (241-249), the
  flipped kernel for backprojection (283-292).

Each case is computed here from those geometric facts in float64 (the
interpolant's integral by the trapezoid rule per kernel segment, which is
exact for a piecewise-linear function) and compared with the oracle's FP64
restatement at 1e-9 and its FP32 restatement at 2e-4.  The cases use
unequal crystal sizes, anisotropic voxels and LORs whose footprint ends land
in the upper half of a voxel, so that swapping the crystal sizes or the
voxel sizes, shifting or flipping the PSF interval, or rounding a footprint
end down instead of to nearest changes some weight by far more than the
tolerance.  The TOF slab's width cancels from its centre in exact
arithmetic, so the TOF case uses a one-voxel-wide minor axis of 2^20 mm:
the slab ends are then rounded to 1/16 mm in FP32, which an exchanged slab
width (the minor voxel size instead of the slice thickness) turns into
weight errors of order 1e-3.
"""
import math

import numpy as np
import pytest

import oracle as O

F64_TOL = 1e-9
F32_TOL = 2e-4
C_MM_PS = 0.299792458


def ovl(a0, a1, b0, b1):
    return max(0.0, min(a1, b1) - max(a0, b0))


def lin_interp_integral(kernel, ks, x0, x1):
    """Exact integral over [x0, x1] of the piecewise-linear function through
    (i - h) * ks -> kernel[i], i = 0..K-1, which falls to 0 at +-(h + 1) * ks
    and is 0 beyond (h = (K - 1) / 2)."""
    k = np.asarray(kernel, np.float64)
    K = k.size
    h = (K - 1) // 2
    vals = np.concatenate([[0.0], k, [0.0]])      # at (j - h - 1) * ks
    xs = (np.arange(K + 2) - h - 1) * ks
    tot = 0.0
    for j in range(K + 1):
        a, b = max(x0, xs[j]), min(x1, xs[j + 1])
        if b <= a:
            continue
        f = lambda t: vals[j] + (vals[j + 1] - vals[j]) * (t - xs[j]) / ks
        tot += 0.5 * (f(a) + f(b)) * (b - a)
    return tot


def centres(n, v):
    return -0.5 * n * v + (np.arange(n) + 0.5) * v


def closed_form(im, p1, p2, tt, tz, major, kernel=None, ks=None, tof=None):
    """Per-voxel DD weights [nz, ny, nx] of one LOR whose endpoints differ
    only along `major` ('x' or 'y') and z, the crystals facing each other
    along `major` (normals -/+ e_major); image offset zero.  kernel: a PSF
    row (the row selected for this LOR); tof: (fwhm_ps, tof_ps), untruncated.
    """
    p1 = np.asarray(p1, np.float64)
    p2 = np.asarray(p2, np.float64)
    ax = 0 if major == "x" else 1
    mi = 1 - ax
    nmaj, vmaj = (im.nx, im.vx) if ax == 0 else (im.ny, im.vy)
    nmin, vmin = (im.ny, im.vy) if ax == 0 else (im.nx, im.vx)
    assert p1[mi] == p2[mi]
    dn = float(np.linalg.norm(p2 - p1))
    w = np.zeros((im.nz, im.ny, im.nx))
    cmaj, cmin, cz = centres(nmaj, vmaj), centres(nmin, vmin), centres(im.nz, im.vz)
    path = vmaj * dn / abs(p2[ax] - p1[ax])     # slice thickness / |cos theta|
    for i, c in enumerate(cmaj):
        a = (c - p1[ax]) / (p2[ax] - p1[ax])
        zc = p1[2] + a * (p2[2] - p1[2])
        f0, f1 = p1[mi] - tt / 2, p1[mi] + tt / 2   # footprint (constant)
        z0, z1 = zc - tz / 2, zc + tz / 2
        wt = 1.0
        if tof is not None:
            fwhm, t_ps = tof
            sig = fwhm * C_MM_PS * 0.5 / (2 * math.sqrt(2 * math.log(2)))
            mu = 0.5 * dn + t_ps * C_MM_PS * 0.5
            wt = math.exp(-0.5 * ((a * dn - mu) / sig) ** 2) / (math.sqrt(2 * math.pi) * sig)
        for j, cm in enumerate(cmin):
            q0, q1 = cm - vmin / 2, cm + vmin / 2
            if kernel is None:
                wm = ovl(q0, q1, f0, f1) / tt
            else:
                wm = lin_interp_integral(kernel, ks, q0 - f1, q1 - f0) / tt
            if wm == 0.0:
                continue
            for k, czk in enumerate(cz):
                wz = ovl(czk - im.vz / 2, czk + im.vz / 2, z0, z1) / tz
                if wz == 0.0:
                    continue
                if ax == 0:
                    w[k, j, i] = wm * wz * path * wt
                else:
                    w[k, i, j] = wm * wz * path * wt
    return w


def one_lor(p1, p2, major, tof=None):
    n = np.zeros(3, np.float32)
    n[0 if major == "x" else 1] = 1.0
    a = lambda v: np.asarray([v], np.float32)
    return O.Lors(a(p1), a(p2), a(-n), a(n), None if tof is None else a(tof))


def check(im, proj, L, ref, back_only=False):
    b64 = O.back(im, proj, L, np.ones(1), dtype=np.float64)[0]
    b32 = O.back(im, proj, L, np.ones(1, np.float32))[0]
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.abs(b64 - ref).max() <= F64_TOL * scale
    assert np.abs(b32 - ref).max() <= F32_TOL * scale
    if not back_only:
        # forward = the same weights against any image (adjoint pair without
        # PSF asymmetry)
        x = np.random.default_rng(11).uniform(0.5, 2.0, (1,) + ref.shape)
        f64 = O.forward(im, x, proj, L, dtype=np.float64)[0]
        assert f64 == pytest.approx(float((ref * x[0]).sum()), rel=F64_TOL)
    return b64


# Unequal crystal sizes and anisotropic voxels throughout.
TT, TZ = float(np.float32(3.2)), 5.0   # the FP32 values the projector holds
IM = O.ImageDef(24, 20, 16, 2.0, 2.5, 3.0)


@pytest.mark.parametrize("major", ["x", "y"])
@pytest.mark.parametrize("zs", [(1.9, 1.9), (-4.6, 7.3), (9.35, -8.2)])
def test_dd_axis_facing_lor_weights(major, zs):
    """Footprints of crystal_size_trans across and crystal_size_z along z;
    the z-tilted cases sweep the footprint ends through every position in a
    voxel (so each rounding of a footprint end is exercised)."""
    R = 180.0
    m = 0.7 if major == "x" else -1.3   # minor-axis offset of the LOR
    if major == "x":
        p1, p2 = [-R, m, zs[0]], [R, m, zs[1]]
    else:
        p1, p2 = [m, -R, zs[0]], [m, R, zs[1]]
    p1, p2 = np.float32(p1), np.float32(p2)
    L = one_lor(p1, p2, major)
    proj = O.ProjDef("dd", crystal_size_z=TZ, crystal_size_trans=TT)
    ref = closed_form(IM, p1, p2, TT, TZ, major)
    check(IM, proj, L, ref)
    # swapping the crystal sizes changes the closed form by far more than
    # the tolerance (the pin above can tell them apart)
    swapped = closed_form(IM, p1, p2, TZ, TT, major)
    assert np.abs(swapped - ref).max() > 100 * F32_TOL * np.abs(ref).max()


def test_dd_footprint_end_rounding():
    """A footprint whose upper z end lies in the upper half of a voxel reaches
    that voxel (the z range is rounded to nearest, ProjectorDD.cpp:515-522):
    here the upper end is at z index 7.6."""
    R = 150.0
    im = O.ImageDef(8, 8, 12, 2.0, 2.0, 2.0)
    # upper end at z index 7.6 (voxel 8 spans 7.5..8.5: 0.2 mm of overlap)
    top = (7.6 - (im.nz - 1) / 2) * im.vz
    zc = top - TZ / 2
    p1, p2 = np.float32([-R, 0.3, zc]), np.float32([R, 0.3, zc])
    L = one_lor(p1, p2, "x")
    proj = O.ProjDef("dd", crystal_size_z=TZ, crystal_size_trans=TT)
    ref = closed_form(im, p1, p2, TT, TZ, "x")
    b = check(im, proj, L, ref)
    assert (ref[8] > 0).any() and np.abs(b[8] - ref[8]).max() <= F64_TOL * np.abs(ref).max()


PSF_K = 0.25   # kernel sample spacing, mm


def psf_proj(kernels, s_step):
    return O.ProjDef("dd", crystal_size_z=TZ, crystal_size_trans=TT,
                     psf_kernels=np.asarray(kernels, np.float32), psf_s_step=s_step,
                     psf_k_spacing=PSF_K)


@pytest.mark.parametrize("major", ["x", "y"])
def test_dd_flat_psf_weights(major):
    """A flat kernel of density c: a voxel well inside the kernel's flat part
    gets c * (voxel width + footprint width) / footprint width across (the
    integral over [p0 - r1, p1 - r0]); near the kernel's ends the linear
    fall-off to the padding sample counts."""
    c = 1.0 / 12.0
    K = 41                       # +-5 mm flat, falling to 0 at +-5.25 mm
    rows = np.stack([np.full(K, c * 3), np.full(K, c), np.full(K, c * 7)]).astype(np.float32)
    c = float(rows[1, 0])
    R = 170.0
    m = 6.1                      # |offset| selects row floor(6.1 / 4) = 1
    if major == "x":
        p1, p2 = [-R, m, -2.0], [R, m, 3.5]
    else:
        p1, p2 = [m, -R, 4.0], [m, R, -1.0]
    p1, p2 = np.float32(p1), np.float32(p2)
    L = one_lor(p1, p2, major)
    proj = psf_proj(rows, 4.0)
    ref = closed_form(IM, p1, p2, TT, TZ, major, kernel=rows[1], ks=PSF_K)
    check(IM, proj, L, ref)
    # a voxel fully inside the flat part: c * (voxel + footprint) / footprint
    vmin = IM.vy if major == "x" else IM.vx
    b = O.back(IM, proj, L, np.ones(1), dtype=np.float64)[0]
    plain = closed_form(IM, p1, p2, TT, TZ, major)
    k0 = np.unravel_index(np.argmax(plain), plain.shape)
    # the most-covered voxel lies inside the flat part: its transaxial factor
    # relative to plain DD's
    ratio = b[k0] / plain[k0]
    cm = centres(IM.ny, IM.vy)[k0[1]] if major == "x" else centres(IM.nx, IM.vx)[k0[2]]
    m = float(p1[1] if major == "x" else p1[0])
    plain_frac = ovl(cm - vmin / 2, cm + vmin / 2, m - TT / 2, m + TT / 2) / TT
    assert ratio * plain_frac == pytest.approx(c * (vmin + TT) / TT, rel=1e-9)


def test_dd_asymmetric_psf_flipped_for_backprojection():
    """An asymmetric kernel: the forward projection integrates the kernel
    row, the backprojection its mirror image (ProjectorDD.cpp:283-292,
    ProjectionPsfManager.cpp:37-45,86-90)."""
    K = 31
    ramp = np.linspace(0.02, 0.3, K)
    rows = np.stack([ramp, ramp[::-1] * 0.5]).astype(np.float32)
    R = 165.0
    p1, p2 = np.float32([-R, 1.1, 0.4]), np.float32([R, 1.1, -3.9])
    L = one_lor(p1, p2, "x")
    proj = psf_proj(rows, 100.0)          # row 0
    fwd_w = closed_form(IM, p1, p2, TT, TZ, "x", kernel=rows[0], ks=PSF_K)
    back_w = closed_form(IM, p1, p2, TT, TZ, "x", kernel=rows[0][::-1], ks=PSF_K)
    assert np.abs(fwd_w - back_w).max() > 0.05 * np.abs(fwd_w).max()
    check(IM, proj, L, back_w, back_only=True)
    x = np.random.default_rng(12).uniform(0.5, 2.0, (1,) + fwd_w.shape)
    f64 = O.forward(IM, x, proj, L, dtype=np.float64)[0]
    f32 = O.forward(IM, x.astype(np.float32), proj, L)[0]
    want = float((fwd_w * x[0]).sum())
    assert f64 == pytest.approx(want, rel=F64_TOL)
    assert f32 == pytest.approx(want, rel=F32_TOL)
    # the mirrored kernel would give a forward value off by far more than
    # the tolerance
    assert abs(float((back_w * x[0]).sum()) - want) > 20 * F32_TOL * want


@pytest.mark.parametrize("tof_ps", [-180.0, 35.0, 120.0])
def test_dd_tof_slab_weights(tof_ps):
    """TOF: the Gaussian at the slab centre's distance from the first
    detector.  The minor axis is one voxel of 2^20 mm, so the slab ends are
    rounded to 1/16 mm in FP32 if the slab width were the minor voxel size."""
    fwhm = 60.0
    big = float(2 ** 20)
    im = O.ImageDef(48, 1, 10, 2.0, big, 3.0)
    R = 190.0
    p1, p2 = np.float32([-R, 0.0, -1.3]), np.float32([R, 0.0, 2.9])
    L = one_lor(p1, p2, "x", tof=tof_ps)
    proj = O.ProjDef("dd", crystal_size_z=TZ, crystal_size_trans=TT,
                     tof_width_ps=fwhm, tof_n_std=0)
    ref = closed_form(im, p1, p2, TT, TZ, "x", tof=(fwhm, tof_ps))
    b64 = O.back(im, proj, L, np.ones(1), dtype=np.float64)[0]
    b32 = O.back(im, proj, L, np.ones(1, np.float32))[0]
    mask = ref > 1e-3 * ref.max()
    assert mask.sum() > 10
    rel64 = np.abs(b64 - ref)[mask] / ref[mask]
    rel32 = np.abs(b32 - ref)[mask] / ref[mask]
    assert rel64.max() <= 1e-9
    assert rel32.max() <= 2e-4
