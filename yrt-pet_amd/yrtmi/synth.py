"""Seeded synthetic phantoms and PET data for tests and benchmarks.

List-mode events are drawn the physical way: an emission point sampled with
probability proportional to the phantom activity (voxel, then uniform within
the voxel), an isotropic direction, the LOR being the pair of crystals the
line hits on the scanner's inner cylinder, and the TOF difference following
the reference's sign convention: the TOF kernel is centred at
d/2 + tof_ps*c/2 from detector 1 (TimeOfFlight.hpp:175-185), so
tof_ps = 2*x/c with x the signed offset of the emission point from the LOR
midpoint toward detector 2, blurred by a Gaussian of the given FWHM.

Everything here has a numpy (CPU) path for small test sizes and a torch
(GPU) path for benchmark sizes; both are seeded.
"""
from __future__ import annotations

import math

import numpy as np

C_MM_PS = 0.299792458


# --------------------------------------------------------------------------
# phantoms
# --------------------------------------------------------------------------
def voxel_centres(im):
    lx, ly, lz = (float(v) for v in im.lengths())
    xs = -lx / 2 + (np.arange(im.nx) + 0.5) * im.vx + im.off_x
    ys = -ly / 2 + (np.arange(im.ny) + 0.5) * im.vy + im.off_y
    zs = -lz / 2 + (np.arange(im.nz) + 0.5) * im.vz + im.off_z
    return xs, ys, zs


def cylinder_spheres(im, radius, length, background, spheres, supersample=2):
    """Water cylinder (axis z) of `radius`/`length` at `background` activity
    plus spheres [(cx, cy, cz, r, activity)] that replace the background,
    averaged over supersample^3 sub-voxel points.  Returns (nz, ny, nx)."""
    xs, ys, zs = voxel_centres(im)
    out = np.zeros((im.nz, im.ny, im.nx), np.float64)
    offs = (np.arange(supersample) + 0.5) / supersample - 0.5
    for ox in offs:
        for oy in offs:
            for oz in offs:
                X = (xs + ox * im.vx)[None, None, :]
                Y = (ys + oy * im.vy)[None, :, None]
                Z = (zs + oz * im.vz)[:, None, None]
                inside = (X ** 2 + Y ** 2 <= radius ** 2) & (np.abs(Z) <= length / 2)
                val = np.where(inside, background, 0.0)
                for (cx, cy, cz, r, a) in spheres:
                    s = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2 <= r ** 2
                    val = np.where(s, a, val)
                out += val
    return (out / supersample ** 3).astype(np.float32)


def random_spheres_in_cylinder(rng, radius, length, radii, activity, margin=0.0):
    """Sphere centres uniform inside the cylinder, each sphere inside it."""
    out = []
    for r in radii:
        while True:
            x, y = rng.uniform(-radius, radius, 2)
            z = rng.uniform(-length / 2, length / 2)
            if math.hypot(x, y) + r + margin <= radius and abs(z) + r + margin <= length / 2:
                out.append((float(x), float(y), float(z), float(r), float(activity)))
                break
    return out


def gaussian_blobs(im, rng, num=10, sigma=(3.0, 15.0), amp=(0.5, 3.0), extent=0.7):
    """Sum of Gaussian blobs with centres uniform in the central `extent` of
    the image."""
    xs, ys, zs = voxel_centres(im)
    lx, ly, lz = (float(v) for v in im.lengths())
    out = np.zeros((im.nz, im.ny, im.nx), np.float64)
    for _ in range(num):
        c = (rng.uniform(-extent, extent) * lx / 2, rng.uniform(-extent, extent) * ly / 2,
             rng.uniform(-extent, extent) * lz / 2)
        s = rng.uniform(*sigma)
        a = rng.uniform(*amp)
        g = (np.exp(-0.5 * ((xs - c[0]) / s) ** 2)[None, None, :]
             * np.exp(-0.5 * ((ys - c[1]) / s) ** 2)[None, :, None]
             * np.exp(-0.5 * ((zs - c[2]) / s) ** 2)[:, None, None])
        out += a * g
    return out.astype(np.float32)


# --------------------------------------------------------------------------
# list-mode events
# --------------------------------------------------------------------------
def random_events(sc, n, rng, tof_fwhm_ps=None, tof_range_ps=600.0):
    """Uniformly random crystal pairs within the maximum ring difference
    (any angle), TOF uniform in +-tof_range_ps: a stress set for parity."""
    dpr, nr = sc.detsPerRing, sc.numRings
    r1 = rng.integers(0, nr, n)
    dz = rng.integers(-sc.maxRingDiff, sc.maxRingDiff + 1, n)
    r2 = np.clip(r1 + dz, 0, nr - 1)
    c1 = rng.integers(0, dpr, n)
    c2 = (c1 + rng.integers(1, dpr, n)) % dpr
    ev = {"d1": (r1 * dpr + c1).astype(np.uint32), "d2": (r2 * dpr + c2).astype(np.uint32)}
    if tof_fwhm_ps is not None:
        ev["tof"] = rng.uniform(-tof_range_ps, tof_range_ps, n).astype(np.float32)
    return ev


def _hit_crystals(xp, p, u, sc):
    """Crystal index hit by the ray p + t u (t > 0) on the inner cylinder of
    radius R; returns (det id, valid)."""
    R = sc.scannerRadius
    a = u[:, 0] ** 2 + u[:, 1] ** 2
    b = 2 * (p[:, 0] * u[:, 0] + p[:, 1] * u[:, 1])
    c = p[:, 0] ** 2 + p[:, 1] ** 2 - R * R
    disc = xp.sqrt(xp.clip(b * b - 4 * a * c, 0, None))
    t = (-b + disc) / (2 * xp.clip(a, 1e-12, None))
    hx = p[:, 0] + t * u[:, 0]
    hy = p[:, 1] + t * u[:, 1]
    hz = p[:, 2] + t * u[:, 2]
    half = sc.axialFOV / 2
    ring = xp.floor((hz + half) / (sc.axialFOV / sc.numRings))
    ang = xp.arctan2(hy, hx) % (2 * math.pi)
    cry = xp.floor(ang / (2 * math.pi / sc.detsPerRing) + 0.5) % sc.detsPerRing
    valid = (ring >= 0) & (ring < sc.numRings) & (a > 1e-12)
    ring = xp.clip(ring, 0, sc.numRings - 1)
    return ring * sc.detsPerRing + cry, valid


def sample_events(sc, im, activity, n, seed, tof_fwhm_ps=None, device=None,
                  batch=1 << 24, motion_xf=None):
    """Draw n detected coincidences from `activity` (nz, ny, nx).  Runs with
    torch on `device` (GPU) when given, else with numpy.  Events outside the
    axial FOV or the ring-difference limit are redrawn.  Returns a dict of
    d1, d2 (uint32 / int32 tensors), tof (float32) and the emission points.

    motion_xf: a rigid transform (12 floats, transform_t rows) that maps the
    scanner coordinates of this stretch of the acquisition to the reference
    position (p_ref = R p + t, as ProjectionData::getLOR applies it): the
    emission points are drawn in the reference position and moved to where
    the object was, p = R^T (p_ref - t), before the detectors are hit."""
    if device is not None:
        import torch
        xp = torch
        g = torch.Generator(device=device)
        g.manual_seed(int(seed))
        act = torch.as_tensor(np.ascontiguousarray(activity), device=device).reshape(-1).double()
        cdf = torch.cumsum(act, 0)
        cdf = cdf / cdf[-1]

        def U(m):
            return torch.rand(m, generator=g, device=device, dtype=torch.float64)

        def N(m):
            return torch.randn(m, generator=g, device=device, dtype=torch.float64)
    else:
        rng = np.random.default_rng(seed)
        xp = np
        act = np.asarray(activity, np.float64).reshape(-1)
        cdf = np.cumsum(act)
        cdf = cdf / cdf[-1]
        U = lambda m: rng.random(m)
        N = lambda m: rng.standard_normal(m)
    lx, ly, lz = (float(v) for v in im.lengths())
    outs = {"d1": [], "d2": [], "tof": []}
    got = 0
    sig_ps = None if tof_fwhm_ps is None else tof_fwhm_ps / (2 * math.sqrt(2 * math.log(2)))
    while got < n:
        m = min(batch, int((n - got) * 1.3) + 1024)
        if device is not None:
            vox = torch.searchsorted(cdf, U(m)).clamp_(max=act.numel() - 1)
        else:
            vox = np.minimum(np.searchsorted(cdf, U(m)), act.size - 1)
        ix = vox % im.nx
        iy = (vox // im.nx) % im.ny
        iz = vox // (im.nx * im.ny)
        px = -lx / 2 + (ix + U(m)) * im.vx + im.off_x
        py = -ly / 2 + (iy + U(m)) * im.vy + im.off_y
        pz = -lz / 2 + (iz + U(m)) * im.vz + im.off_z
        cz = 2 * U(m) - 1
        ph = 2 * math.pi * U(m)
        sz = xp.sqrt(1 - cz * cz)
        u = xp.stack([sz * xp.cos(ph), sz * xp.sin(ph), cz], 1)
        p = xp.stack([px, py, pz], 1)
        if motion_xf is not None:
            M = np.asarray(motion_xf, np.float64).reshape(3, 4)
            Rt = M[:, :3].T
            if device is not None:
                Rt = torch.as_tensor(Rt, device=device)
                t = torch.as_tensor(M[:, 3], device=device)
            else:
                t = M[:, 3]
            p = (p - t) @ Rt.T
        da, va = _hit_crystals(xp, p, u, sc)
        db, vb = _hit_crystals(xp, p, -u, sc)
        ra, rb = da // sc.detsPerRing, db // sc.detsPerRing
        ok = va & vb & (xp.abs(ra - rb) <= sc.maxRingDiff) & (da != db)
        swap = U(m) < 0.5
        d1 = xp.where(swap, db, da)
        d2 = xp.where(swap, da, db)
        if sig_ps is not None:
            # crystal-centre LUT endpoints; x = projection of the emission
            # point on the LOR, from the midpoint toward d2
            pos = _lut_xp(sc, xp, device)
            q1 = pos[d1.long() if device is not None else d1.astype(np.int64)]
            q2 = pos[d2.long() if device is not None else d2.astype(np.int64)]
            dv = q2 - q1
            dn = xp.sqrt((dv * dv).sum(1))
            mid = 0.5 * (q1 + q2)
            xoff = ((p - mid) * dv).sum(1) / dn
            tof = 2 * xoff / C_MM_PS + sig_ps * N(m)
        if device is not None:
            sel = ok.nonzero().squeeze(1)
        else:
            sel = np.nonzero(ok)[0]
        take = min(int(sel.shape[0]), n - got)
        sel = sel[:take]
        outs["d1"].append(d1[sel])
        outs["d2"].append(d2[sel])
        if sig_ps is not None:
            outs["tof"].append(tof[sel])
        got += take
    if device is not None:
        import torch
        res = {"d1": torch.cat(outs["d1"]).to(torch.int32),
               "d2": torch.cat(outs["d2"]).to(torch.int32)}
        if sig_ps is not None:
            res["tof"] = torch.cat(outs["tof"]).float()
    else:
        res = {"d1": np.concatenate(outs["d1"]).astype(np.uint32),
               "d2": np.concatenate(outs["d2"]).astype(np.uint32)}
        if sig_ps is not None:
            res["tof"] = np.concatenate(outs["tof"]).astype(np.float32)
    return res


_LUT_CACHE = {}


def _lut_xp(sc, xp, device):
    key = (id(sc), str(device))
    if key not in _LUT_CACHE:
        pos, _ = sc.lut()
        if device is not None:
            import torch
            _LUT_CACHE[key] = torch.as_tensor(pos, device=device).double()
        else:
            _LUT_CACHE[key] = pos.astype(np.float64)
    return _LUT_CACHE[key]


def rigid_transforms(rng, k, max_rot_deg=3.0, max_trans_mm=5.0):
    """k seeded rigid transforms as transform_t rows (r00 r01 r02 tx r10 ...;
    include/yrt-pet/utils/Types.hpp:57-60)."""
    out = np.zeros((k, 12), np.float32)
    for i in range(k):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        ang = math.radians(rng.uniform(-max_rot_deg, max_rot_deg))
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K
        t = rng.normal(size=3)
        t *= rng.uniform(0, max_trans_mm) / np.linalg.norm(t)  # |t| <= max_trans_mm
        out[i] = np.array([R[0, 0], R[0, 1], R[0, 2], t[0], R[1, 0], R[1, 1], R[1, 2], t[1],
                           R[2, 0], R[2, 1], R[2, 2], t[2]], np.float32)
    return out


def gaussian_psf_row(fwhm_mm, k_spacing=0.25, num_s=1, s_step=1000.0, nsig=4.0):
    """Projection-space PSF in the one-row CSV layout of
    ProjectionPsfManager (sStep, kSpacing, K, K*numS taps): a normalised
    Gaussian density sampled at odd K taps."""
    sig = fwhm_mm / (2 * math.sqrt(2 * math.log(2)))
    half = int(math.ceil(nsig * sig / k_spacing))
    xs = (np.arange(-half, half + 1)) * k_spacing
    k = np.exp(-0.5 * (xs / sig) ** 2) / (math.sqrt(2 * math.pi) * sig)
    ker = np.tile(k.astype(np.float32), (num_s, 1))
    return ker, s_step, k_spacing
