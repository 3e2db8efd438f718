"""TEST INFRASTRUCTURE ONLY - CPU oracle for the OSEM projection path.

ctypes front end of ``liboracle.so`` (the C++ restatement in
``yrt_oracle.cpp``) plus the OSEM driver flow restated in numpy.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker (or as the CPU
baseline being timed); the product package never imports it.

Reference files are cited relative to the YRT-PET 2.0.7 source tree.

Parity status of each function is recorded in its docstring ("pinned by").
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

EPS_FLT = np.float32(1.19209e-07)  # geometry/Constants.hpp:28
INITIAL_VALUE_MLEM = np.float32(0.125)  # recon/OSEM.hpp:36
DEFAULT_HARD_THRESHOLD = np.float32(1.0)  # recon/OSEM.hpp:34
DEFAULT_DENOM_THRESHOLD = np.float32(1e-35)  # recon/OSEM.hpp:35


class _Image(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("nt", C.c_int64), ("vx", C.c_float), ("vy", C.c_float),
                ("vz", C.c_float), ("lx", C.c_float), ("ly", C.c_float),
                ("lz", C.c_float), ("ox", C.c_float), ("oy", C.c_float),
                ("oz", C.c_float), ("fov_radius", C.c_float)]


class _Scanner(C.Structure):
    _fields_ = [("axial_fov", C.c_float), ("crystal_size_z", C.c_float),
                ("crystal_size_trans", C.c_float),
                ("crystal_depth", C.c_float), ("radius", C.c_float),
                ("dets_per_ring", C.c_int64), ("num_rings", C.c_int64),
                ("num_doi", C.c_int64), ("max_ring_diff", C.c_int64),
                ("min_ang_diff", C.c_int64), ("dets_per_block", C.c_int64)]


class _Projector(C.Structure):
    _fields_ = [("type", C.c_int), ("num_rays", C.c_int),
                ("use_tof", C.c_int), ("tof_width_ps", C.c_float),
                ("tof_n_std", C.c_int), ("ray_offsets", C.c_void_p),
                ("psf_kernels", C.c_void_p), ("psf_num_s", C.c_int),
                ("psf_ksize", C.c_int), ("psf_s_step", C.c_float),
                ("psf_k_spacing", C.c_float), ("crystal_size_z", C.c_float),
                ("crystal_size_trans", C.c_float),
                ("crystal_depth", C.c_float)]


class _Lors(C.Structure):
    _fields_ = [("n", C.c_int64), ("p1", C.c_void_p), ("p2", C.c_void_p),
                ("n1", C.c_void_p), ("n2", C.c_void_p), ("tof", C.c_void_p),
                ("frame", C.c_void_p)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        _LIB = C.CDLL(path)
        _LIB.yo_tof_weight.restype = C.c_float
        _LIB.yo_psf_integral.restype = C.c_float
        _LIB.yo_dd_overlap.restype = C.c_float
        _LIB.yo_histo3d_subset_size.restype = C.c_int64
        for f in ("yo_tof_weight",):
            getattr(_LIB, f).argtypes = [C.c_float, C.c_int, C.c_float,
                                         C.c_float, C.c_float, C.c_float]
        _LIB.yo_psf_integral.argtypes = [C.c_void_p, C.c_int, C.c_float,
                                         C.c_float, C.c_float]
        _LIB.yo_dd_overlap.argtypes = [C.c_float] * 4
        _LIB.yo_get_alpha.argtypes = [C.c_float] * 5 + [C.c_void_p] * 2
        _LIB.yo_tof_alpha_range.argtypes = [C.c_float, C.c_int, C.c_float,
                                            C.c_float, C.c_void_p, C.c_void_p]
        _LIB.yo_tof_helper.argtypes = [C.c_float, C.c_int] + [C.c_void_p] * 3
        _LIB.yo_em_subset_f32.argtypes = [C.c_void_p] * 11 + [C.c_float, C.c_float, C.c_int]
        _LIB.yo_em_subset_f64.argtypes = [C.c_void_p] * 11 + [C.c_double, C.c_double, C.c_int]
        _LIB.yo_em_update_f32.argtypes = [C.c_void_p] * 3 + [C.c_int64, C.c_float]
        _LIB.yo_em_update_f64.argtypes = [C.c_void_p] * 3 + [C.c_int64, C.c_double]
        _LIB.yo_apply_threshold_f32.argtypes = [C.c_void_p] * 2 + [C.c_int64] + [C.c_float] * 5
    return _LIB


def ref_lib():
    """The reference's own TOF / get_alpha / PSF / multi-ray code, compiled
    from the reference tree into oracle/_ref (None when not built)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libyrtref.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
        _REF.yr_tof_weight.restype = C.c_float
        _REF.yr_tof_weight.argtypes = [C.c_float, C.c_int] + [C.c_float] * 4
        _REF.yr_tof_alpha_range.argtypes = [C.c_float, C.c_int, C.c_float,
                                            C.c_float, C.c_void_p, C.c_void_p]
        _REF.yr_tof_helper.argtypes = [C.c_float, C.c_int] + [C.c_void_p] * 3
        _REF.yr_get_alpha.argtypes = [C.c_float] * 5 + [C.c_void_p] * 2
        _REF.yr_calc_integral.restype = C.c_float
        _REF.yr_calc_integral.argtypes = [C.c_void_p, C.c_int, C.c_float,
                                          C.c_float, C.c_float]
        _REF.yr_multiray_lines.argtypes = [C.c_void_p] * 4 + [C.c_float] * 3 + [C.c_int, C.c_void_p]
        _REF.yr_psf_query.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p,
                                      C.c_int, C.c_float, C.c_float,
                                      C.c_void_p, C.c_void_p, C.c_void_p]
    return _REF


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------
# geometry descriptors
# --------------------------------------------------------------------------
@dataclass
class ScannerDef:
    """Regular cylindrical scanner (Scanner ctor, src/datastruct/scanner/
    Scanner.cpp:101-127)."""
    axial_fov: float
    crystal_size_z: float
    crystal_size_trans: float
    crystal_depth: float
    radius: float
    dets_per_ring: int
    num_rings: int
    num_doi: int = 1
    max_ring_diff: int = -1
    min_ang_diff: int = 1
    dets_per_block: int = 1

    def __post_init__(self):
        if self.max_ring_diff < 0:
            self.max_ring_diff = self.num_rings - 1

    @property
    def num_dets(self):
        return self.dets_per_ring * self.num_rings * self.num_doi

    def c(self):
        return _Scanner(self.axial_fov, self.crystal_size_z,
                        self.crystal_size_trans, self.crystal_depth,
                        self.radius, self.dets_per_ring, self.num_rings,
                        self.num_doi, self.max_ring_diff, self.min_ang_diff,
                        self.dets_per_block)


@dataclass
class ImageDef:
    """ImageParams (src/datastruct/image/ImageParams.cpp:136-155,206-262)."""
    nx: int
    ny: int
    nz: int
    vx: float
    vy: float
    vz: float
    ox: float = 0.0
    oy: float = 0.0
    oz: float = 0.0
    nt: int = 1
    lx: float | None = None  # explicit lengths (ImageParams built from lengths)
    ly: float | None = None
    lz: float | None = None

    @staticmethod
    def from_lengths(nx, ny, nz, lx, ly, lz, ox=0.0, oy=0.0, oz=0.0, nt=1):
        """ImageParams(nx, ny, nz, length_x, ...): v = length / n in float32,
        the lengths kept as given (ImageParams.cpp:256-262)."""
        f = np.float32
        return ImageDef(nx, ny, nz, float(f(lx) / f(nx)), float(f(ly) / f(ny)),
                        float(f(lz) / f(nz)), ox, oy, oz, nt, float(f(lx)),
                        float(f(ly)), float(f(lz)))

    @property
    def lengths(self):
        """length = v * n when built from voxel sizes (ImageParams.cpp:263-266)."""
        f = np.float32
        return (f(self.lx) if self.lx is not None else f(self.vx) * f(self.nx),
                f(self.ly) if self.ly is not None else f(self.vy) * f(self.ny),
                f(self.lz) if self.lz is not None else f(self.vz) * f(self.nz))

    @property
    def fov_radius(self):
        lx, ly, _ = self.lengths
        r = max(lx / np.float32(2), ly / np.float32(2))
        return np.float32(r - max(np.float32(self.vx), np.float32(self.vy)) / np.float32(1000))

    @property
    def shape(self):
        return (self.nt, self.nz, self.ny, self.nx)

    @property
    def nvox(self):
        return self.nx * self.ny * self.nz

    def c(self, zero_offset=False):
        lx, ly, lz = self.lengths
        o = (0.0, 0.0, 0.0) if zero_offset else (self.ox, self.oy, self.oz)
        return _Image(self.nx, self.ny, self.nz, self.nt, self.vx, self.vy,
                      self.vz, lx, ly, lz, o[0], o[1], o[2], self.fov_radius)


@dataclass
class ProjDef:
    """ProjectorParams (include/yrt-pet/operators/ProjectorParams.hpp:27-71)
    plus the projection-space PSF rows (ProjectionPsfManager)."""
    type: str = "siddon"  # or "dd"
    num_rays: int = 1
    tof_width_ps: float | None = None
    tof_n_std: int = 0
    psf_kernels: np.ndarray | None = None  # (num_s, K) float32
    psf_s_step: float = 0.0
    psf_k_spacing: float = 0.0
    crystal_size_z: float = 0.0
    crystal_size_trans: float = 0.0
    crystal_depth: float = 0.0
    _keep: list = field(default_factory=list, repr=False)

    def without_tof(self):
        return ProjDef(self.type, self.num_rays, None, 0, self.psf_kernels,
                       self.psf_s_step, self.psf_k_spacing,
                       self.crystal_size_z, self.crystal_size_trans,
                       self.crystal_depth)

    def c(self):
        offs = multiray_offsets(self.num_rays) if self.num_rays > 1 else None
        ker = None
        if self.psf_kernels is not None:
            ker = np.ascontiguousarray(self.psf_kernels, dtype=np.float32)
        self._keep = [offs, ker]
        return _Projector(
            1 if self.type.lower() == "dd" else 0, self.num_rays,
            1 if self.tof_width_ps is not None else 0,
            float(self.tof_width_ps or 0.0), int(self.tof_n_std),
            _p(offs), _p(ker),
            0 if ker is None else ker.shape[0],
            0 if ker is None else ker.shape[1], self.psf_s_step,
            self.psf_k_spacing, self.crystal_size_z, self.crystal_size_trans,
            self.crystal_depth)


def proj_for_scanner(sc: ScannerDef, **kw) -> ProjDef:
    return ProjDef(crystal_size_z=sc.crystal_size_z,
                   crystal_size_trans=sc.crystal_size_trans,
                   crystal_depth=sc.crystal_depth, **kw)


# --------------------------------------------------------------------------
# data model
# --------------------------------------------------------------------------
def lut(sc: ScannerDef):
    """DetRegular::generateLUT (src/datastruct/scanner/DetRegular.cpp:113-182).
    Pinned by: closed-form ring geometry (tests/test_oracle_geometry.py)."""
    n = sc.num_dets
    pos = np.zeros((n, 3), np.float32)
    ori = np.zeros((n, 3), np.float32)
    s = sc.c()
    lib().yo_lut_regular(C.byref(s), _p(pos), _p(ori))
    return pos, ori


def histo_dims(sc: ScannerDef):
    """Histogram3D ctor dims (src/datastruct/projection/Histogram3D.cpp:
    165-203) -> (numZBin, numPhi, numR)."""
    r, p, z = C.c_int64(), C.c_int64(), C.c_int64()
    s = sc.c()
    lib().yo_histo3d_dims(C.byref(s), C.byref(r), C.byref(p), C.byref(z))
    return z.value, p.value, r.value


def histo_pairs(sc: ScannerDef, bins):
    """Histogram3D::getDetPairFromCoords (Histogram3D.cpp:299-363).
    Pinned by: bin<->pair round trip, uniqueness, ring-difference rule."""
    bins = np.ascontiguousarray(bins, np.int64)
    d1 = np.empty(bins.size, np.uint32)
    d2 = np.empty(bins.size, np.uint32)
    s = sc.c()
    lib().yo_histo3d_pairs(C.byref(s), _p(bins), C.c_int64(bins.size), _p(d1), _p(d2))
    return d1, d2


def histo_bins_from_pairs(sc: ScannerDef, d1, d2):
    """Histogram3D::getBinIdFromDetPair (Histogram3D.cpp:366-421)."""
    d1 = np.ascontiguousarray(d1, np.uint32)
    d2 = np.ascontiguousarray(d2, np.uint32)
    out = np.empty(d1.size, np.int64)
    s = sc.c()
    lib().yo_histo3d_bins_from_pairs(C.byref(s), _p(d1), _p(d2), C.c_int64(d1.size), _p(out))
    return out


def histo_subset_bins(sc: ScannerDef, num_subsets, subset):
    """BinIteratorRangeHistogram3D (src/datastruct/projection/BinIterator.cpp:
    147-198).  Pinned by: the subsets partition all bins."""
    s = sc.c()
    n = lib().yo_histo3d_subset_size(C.byref(s), int(num_subsets), int(subset))
    out = np.empty(n, np.int64)
    lib().yo_histo3d_subset_bins(C.byref(s), int(num_subsets), int(subset), _p(out))
    return out


def listmode_subset(num_events, num_subsets, subset):
    """BinIteratorChronologicalInterleaved (BinIterator.cpp:261-265):
    events subset, subset+S, ...  (BinIteratorRange size formula 111-114)."""
    return np.arange(subset, num_events, num_subsets, dtype=np.int64)


def multiray_offsets(num_rays):
    """The 4 rand_r(seed=13) uniforms per extra ray
    (src/geometry/MultiRayGenerator.cpp:78-92; ProjectorSiddon.cpp:258,330).
    Pinned by: the reference MultiRayGenerator compiled in oracle/_ref."""
    out = np.zeros(max(0, 4 * (num_rays - 1)), np.float32)
    if num_rays > 1:
        lib().yo_multiray_offsets(int(num_rays), _p(out))
    return out


def multiray_lines(p1, p2, n1, n2, thick_z, thick_trans, num_lines):
    """Rays 1..num_lines of MultiRayGenerator::getRandomLine for one LOR
    (seed 13, not parallel, no depth).  Pinned by: the reference generator
    compiled in oracle/_ref (tests/golden/ref_vectors.npz)."""
    f = lambda a: np.ascontiguousarray(a, np.float32)
    out = np.zeros((num_lines, 6), np.float32)
    lib().yo_multiray_lines(_p(f(p1)), _p(f(p2)), _p(f(n1)), _p(f(n2)),
                            C.c_float(thick_z), C.c_float(thick_trans),
                            int(num_lines), _p(out))
    return out


def apply_motion(p, xf):
    """ProjectionData::getLOR motion branch (src/datastruct/projection/
    ProjectionData.cpp:302-318) with Matrix*Vector3D (src/geometry/Matrix.cpp:
    177-186), in float32.  p: (n,3); xf: (n,12) rows of transform_t."""
    f = np.float32
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    r = xf.astype(f)
    ox = (r[:, 0] * x + r[:, 1] * y) + r[:, 2] * z
    oy = (r[:, 4] * x + r[:, 5] * y) + r[:, 6] * z
    oz = (r[:, 8] * x + r[:, 9] * y) + r[:, 10] * z
    return np.stack([ox + r[:, 3], oy + r[:, 7], oz + r[:, 11]], axis=1).astype(f)


@dataclass
class Lors:
    p1: np.ndarray
    p2: np.ndarray
    n1: np.ndarray | None = None
    n2: np.ndarray | None = None
    tof: np.ndarray | None = None
    frame: np.ndarray | None = None

    @property
    def n(self):
        return self.p1.shape[0]

    def take(self, idx):
        g = (lambda a: None if a is None else np.ascontiguousarray(a[idx]))
        return Lors(g(self.p1), g(self.p2), g(self.n1), g(self.n2), g(self.tof), g(self.frame))

    def c(self):
        for name in ("p1", "p2", "n1", "n2", "tof"):
            a = getattr(self, name)
            if a is not None:
                setattr(self, name, np.ascontiguousarray(a, np.float32))
        if self.frame is not None:
            self.frame = np.ascontiguousarray(self.frame, np.int32)
        return _Lors(self.n, _p(self.p1), _p(self.p2), _p(self.n1),
                     _p(self.n2), _p(self.tof), _p(self.frame))


def lors_from_pairs(pos, ori, d1, d2, tof=None, frame=None, motion_frame=None,
                    motion_xf=None):
    """ProjectionData::collectProjectionProperties / getLOR
    (src/datastruct/projection/ProjectionData.cpp:198-320): endpoints from the
    detector LUT (motion applied to endpoints only), orientations from the
    LUT."""
    d1 = np.asarray(d1, np.int64)
    d2 = np.asarray(d2, np.int64)
    p1 = pos[d1]
    p2 = pos[d2]
    if motion_frame is not None:
        xf = np.asarray(motion_xf, np.float32)[np.asarray(motion_frame, np.int64)]
        p1 = apply_motion(p1, xf)
        p2 = apply_motion(p2, xf)
    return Lors(p1, p2, ori[d1], ori[d2],
                None if tof is None else np.asarray(tof, np.float32),
                None if frame is None else np.asarray(frame, np.int32))


# --------------------------------------------------------------------------
# helpers exposed for pinning
# --------------------------------------------------------------------------
def tof_helper(width_ps, n_std):
    a, b, c = C.c_float(), C.c_float(), C.c_float()
    lib().yo_tof_helper(width_ps, n_std, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def tof_alpha_range(width_ps, n_std, lor_norm, tof_ps):
    a, b = C.c_float(), C.c_float()
    lib().yo_tof_alpha_range(width_ps, n_std, lor_norm, tof_ps, C.byref(a), C.byref(b))
    return a.value, b.value


def tof_weight(width_ps, n_std, lor_norm, tof_ps, lo, hi):
    return lib().yo_tof_weight(width_ps, n_std, lor_norm, tof_ps, lo, hi)


def psf_integral(kernel, kstep, x0, x1):
    k = np.ascontiguousarray(kernel, np.float32)
    return lib().yo_psf_integral(_p(k), k.size, kstep, x0, x1)


def dd_overlap(p0, p1, d0, d1):
    return lib().yo_dd_overlap(p0, p1, d0, d1)


def get_alpha(r0, r1, p1, p2, inv):
    a, b = C.c_float(), C.c_float()
    lib().yo_get_alpha(r0, r1, p1, p2, inv, C.byref(a), C.byref(b))
    return a.value, b.value


# --------------------------------------------------------------------------
# projectors
# --------------------------------------------------------------------------
def _dt(dtype):
    return np.float64 if np.dtype(dtype) == np.float64 else np.float32


def forward(im: ImageDef, img, proj: ProjDef, lors: Lors, dtype=np.float32,
            nthreads=1, zero_offset=False):
    """Forward projection of each LOR: ProjectorSiddon / ProjectorDD
    forwardProjection (src/operators/ProjectorSiddon.cpp:239-307,
    src/operators/ProjectorDD.cpp:154-192)."""
    dt = _dt(dtype)
    img = np.ascontiguousarray(img, dt)
    out = np.zeros(lors.n, dt)
    ci, cp, cl = im.c(zero_offset), proj.c(), lors.c()
    f = lib().yo_fwd_f64 if dt == np.float64 else lib().yo_fwd_f32
    f(C.byref(ci), _p(img), C.byref(cp), C.byref(cl), _p(out), C.c_int(nthreads))
    return out


def back(im: ImageDef, proj: ProjDef, lors: Lors, vals, dtype=np.float32,
         nthreads=1, out=None):
    """Backprojection (ProjectorSiddon.cpp:309-369, ProjectorDD.cpp:194-232),
    accumulating into `out` (a zero image if None)."""
    dt = _dt(dtype)
    vals = np.ascontiguousarray(vals, dt)
    if out is None:
        out = np.zeros(im.shape, dt)
    ci, cp, cl = im.c(), proj.c(), lors.c()
    f = lib().yo_back_f64 if dt == np.float64 else lib().yo_back_f32
    f(C.byref(ci), _p(out), C.byref(cp), C.byref(cl), _p(vals), C.c_int(nthreads))
    return out


def count_segments(im: ImageDef, proj: ProjDef, lors: Lors, nthreads=1):
    """Voxel segments (Siddon) / footprint voxels (DD) touched per LOR, summed
    over rays: the algorithmic work unit."""
    out = np.zeros(lors.n, np.int64)
    ci, cp, cl = im.c(), proj.c(), lors.c()
    lib().yo_count_segments(C.byref(ci), C.byref(cp), C.byref(cl), _p(out), C.c_int(nthreads))
    return out


def em_subset(im, x, u, proj, lors, meas=None, sens=None, att=None, rnd=None,
              scat=None, global_scale=1.0, denom_thr=DEFAULT_DENOM_THRESHOLD,
              dtype=np.float32, nthreads=1, precorr=None):
    """OSEM_CPU::computeEMUpdateImage (src/recon/OSEM_CPU.cpp:284-393),
    accumulating A^T(m / (A x)) into u (in place).  precorr: the in-vivo
    attenuation precorrection, multiplying the denominator after the
    additive terms (OSEM_GPU.cu:1303-1308).  Pinned by: precorr p equals
    measurements m / p (tests/test_oracle_attenuation.py)."""
    dt = _dt(dtype)
    assert x.dtype == dt and u.dtype == dt and u.flags.c_contiguous
    cv = lambda a: None if a is None else np.ascontiguousarray(a, dt)
    meas, sens, att, rnd, scat, precorr = map(cv, (meas, sens, att, rnd, scat, precorr))
    ci, cp, cl = im.c(), proj.c(), lors.c()
    f = lib().yo_em_subset_f64 if dt == np.float64 else lib().yo_em_subset_f32
    f(C.byref(ci), _p(x), _p(u), C.byref(cp), C.byref(cl), _p(meas),
      _p(sens), _p(att), _p(rnd), _p(scat), _p(precorr), global_scale, denom_thr, nthreads)
    return u


def em_update(x, u, s, thr=EPS_FLT):
    """Image::updateEMThreshold{Static,DynamicWithScaling (all ones)}
    (src/datastruct/image/Image.cpp:754-833): x *= u / s where s > thr, the 3D
    sensitivity broadcast over frames."""
    dt = x.dtype
    nvox = s.size
    xf = x.reshape(-1)
    uf = np.ascontiguousarray(u, dt).reshape(-1)
    sf = np.ascontiguousarray(s, dt).reshape(-1)
    f = lib().yo_em_update_f64 if dt == np.float64 else lib().yo_em_update_f32
    for t in range(xf.size // nvox):
        xs = xf[t * nvox:(t + 1) * nvox]
        us = np.ascontiguousarray(uf[t * nvox:(t + 1) * nvox])
        f(_p(xs), _p(us), _p(sf), C.c_int64(nvox), thr)
    return x


def apply_threshold(x, mask, thr, le_scale, le_off, gt_scale, gt_off):
    """Image::applyThreshold (src/datastruct/image/Image.cpp:595-618)."""
    x = np.ascontiguousarray(x, np.float32)
    m = np.ascontiguousarray(mask, np.float32)
    lib().yo_apply_threshold_f32(_p(x), _p(m), C.c_int64(x.size), thr,
                                 le_scale, le_off, gt_scale, gt_off)
    return x


def attenuation_factors(att_im: ImageDef, att_img, lors: Lors, nthreads=1):
    """Corrector::getAttenuationFactorFromAttenuationImage
    (src/recon/Corrector.cpp:601-611): single-ray Siddon of the mu image
    (ProjectorSiddon::singleForwardProjection, which does not subtract the
    image offset) then exp(-0.1 * p) (utils/Tools.hpp:98-102)."""
    p = forward(att_im, att_img, ProjDef("siddon", 1), Lors(lors.p1, lors.p2),
                nthreads=nthreads, zero_offset=True)
    return np.exp(-p * np.float32(0.1)).astype(np.float32)


# --------------------------------------------------------------------------
# OSEM driver (restated)
# --------------------------------------------------------------------------
def psf_convolve(img, kx, ky, kz, transpose=False):
    """OperatorPsf::convolve (src/operators/OperatorPsf.cpp:290-375): per
    frame, the separable convolution along x, then y, then z, each output
    out[i] = sum over kk = -h..h of k[kk + h] * in[circular(i - kk)] in FP32,
    kk ascending (circular: i - kk wrapped once, Tools.cpp:160-167).
    transpose: the flipped kernels (applyAH, OperatorPsf.cpp:268-276).
    np.roll(a, kk)[i] = a[(i - kk) mod n], which is circular(n, i - kk) for
    |kk| < n.  Pinned by: the identity kernel, a brute-force loop on small
    images and the adjoint identity (tests/test_oracle_psf.py)."""
    x = np.asarray(img, np.float32)
    shape = x.shape
    x = x.reshape((-1,) + shape[-3:])  # frames, nz, ny, nx
    ks = [np.asarray(k, np.float32) for k in (kx, ky, kz)]
    if transpose:
        ks = [k[::-1].copy() for k in ks]
    for k, axis in zip(ks, (3, 2, 1)):
        h = len(k) // 2
        if len(k) % 2 == 0 or h >= x.shape[axis]:
            raise ValueError("kernel length must be odd and its half below the image size")
        acc = np.zeros_like(x)
        for kk in range(-h, h + 1):
            acc = acc + k[kk + h] * np.roll(x, kk, axis=axis)
        x = acc.astype(np.float32)
    return x.reshape(shape)


def gaussian_kernel_1d(sigma, voxel_size, size=0):
    """OperatorPsf::createGaussianKernel1D (OperatorPsf.cpp:158-185): size
    (0: rint(5 sigma / v)) taps exp(-0.5 ((i - size/2) v / sigma)^2), FP32,
    normalised by their FP32 sum."""
    sigma, v = np.float32(sigma), np.float32(voxel_size)
    if size == 0:
        size = int(np.rint(np.float32(5) * sigma / v))
    x = (np.arange(size) - size // 2).astype(np.float32)
    k = np.exp(np.float32(-0.5) * ((x * v / sigma) ** 2)).astype(np.float32)
    tot = np.float32(0)
    for t in k:
        tot = np.float32(tot + t)
    return (k / tot).astype(np.float32)


# --------------------------------------------------------------------------
# Motion: time-averaged (moved) images
# --------------------------------------------------------------------------
def _fmaf(a, b, c):
    """fmaf in float32 through float64: a*b of two floats is exact in
    float64, the sum is rounded to float64 and then to float32 (a double
    rounding that can differ from one true fmaf by one ulp, rarely)."""
    return (np.float64(1.0) * np.asarray(a, np.float32) * np.asarray(b, np.float32)
            + np.asarray(c, np.float32)).astype(np.float32)


def invert_transform(xf):
    """util::invertTransform (include/yrt-pet/geometry/TransformUtils.hpp:
    16-40): rigid inverse [R^T | -R^T t] of transform_t rows (r00 r01 r02 tx
    r10 ... tz), in FP32, each dot product summed left to right.
    Pinned by: composition with the forward transform is the identity on
    points (tests/test_oracle_motion.py)."""
    f = np.float32
    a = np.asarray(xf, f).reshape(-1, 12)
    R = a[:, [0, 1, 2, 4, 5, 6, 8, 9, 10]].reshape(-1, 3, 3)
    t = a[:, [3, 7, 11]]
    RT = np.transpose(R, (0, 2, 1))
    out = np.empty_like(a)
    for i in range(3):
        ti = -((RT[:, i, 0] * t[:, 0] + RT[:, i, 1] * t[:, 1]) + RT[:, i, 2] * t[:, 2])
        out[:, 4 * i:4 * i + 3] = RT[:, i, :]
        out[:, 4 * i + 3] = ti.astype(f)
    return out.reshape(np.shape(xf))


def full_time_range(timestamps):
    """util::getFullTimeRange (src/utils/ReconstructionUtils.cpp:468-495):
    [first start, first start + 1] for one frame, else [first start, last
    start + the second-to-last frame's duration]."""
    t = [int(v) for v in timestamps]
    if not t:
        raise ValueError("no motion frames")
    if len(t) == 1:
        return t[0], t[0] + 1
    return t[0], t[-1] + (t[-1] - t[-2])


def motion_duration(timestamps, frame):
    """LORMotion::getDuration (src/datastruct/projection/LORMotion.cpp:
    73-88): the next frame's start minus this one's; the last frame takes
    the second-to-last frame's duration.  As float."""
    t = [int(v) for v in timestamps]
    last = len(t) - 1
    if frame < last:
        return np.float32(t[frame + 1] - t[frame])
    if frame == last:
        return np.float32(t[last] - t[last - 1])
    raise IndexError("Frame index out of range")


def time_average_frames(timestamps, xf, t0, t1):
    """The frames and weights of a time average over [t0, t1]
    (timeAverageMoveImageDevice, src/utils/ReconstructionUtilsDevice.cu:
    146-190): the frames whose start lies in [t0, t1] (both ends included),
    each weighted duration / float(t1 - t0), with its inverted transform.
    Returns (frame indices, inverse transforms (k,12), weights (k,))."""
    if not t1 > t0:
        raise ValueError("Time stop must be greater than time start")
    t = [int(v) for v in timestamps]
    frames = [f for f in range(len(t)) if t0 <= t[f] <= t1]
    # the loop stops at the first start beyond t1; starts are increasing
    stop = next((f for f in range(len(t)) if t[f] >= t0 and t[f] > t1), len(t))
    frames = [f for f in frames if f < stop]
    if not frames:
        raise ValueError("no motion frame starts inside the time range")
    scan = np.float32(t1 - t0)
    if len(t) == 1:
        # one motion frame: LORMotion::getDuration would read before the
        # first record; the one frame covers the whole range
        w = np.ones(len(frames), np.float32)
    else:
        w = np.array([motion_duration(t, f) / scan for f in frames], np.float32)
    inv = invert_transform(np.asarray(xf, np.float32).reshape(-1, 12)[frames])
    return np.array(frames), inv, w


def _image_geometry(im: ImageDef):
    f = np.float32
    n = (im.nx, im.ny, im.nz)
    L = im.lengths
    off = (f(im.ox), f(im.oy), f(im.oz))
    v = tuple(f(L[a] / f(n[a])) for a in range(3))
    inv = tuple(f(f(1) / v[a]) for a in range(3))
    origin = tuple(f(f(off[a] - f(L[a] * f(0.5))) + f(v[a] * f(0.5))) for a in range(3))
    # util::indexToPosition (ImageUtils.cuh:96-101): ((i v - L/2) + off) + v/2
    pos = tuple(((np.arange(n[a], dtype=f) * v[a] - f(f(0.5) * L[a])) + off[a])
                + f(f(0.5) * v[a]) for a in range(3))
    return n, inv, origin, pos


def move_accumulate(img, im: ImageDef, inv_xf, weights, out):
    """timeAverageMoveImage_kernel<WEIGHED_AVG=true> (src/datastruct/image/
    ImageSpaceKernels.cu:191-264) with util::trilinearInterpolateCore
    (include/yrt-pet/datastruct/image/ImageUtils.cuh:15-70): for every
    output voxel centre p and every frame, q = inv * p (nested fmaf as the
    kernel), c = (q - origin) * inv_v, i0 = floor(c), d = c - i0, the 8
    corner weights (1-dx)(1-dy)(1-dz) ... (corners outside the image get
    weight 0), and out += in[corner] * w * frame_weight (one fmaf), corners
    in the kernel's order, frames in order.  img: (nz, ny, nx) frame 0 of
    the input; out: (nz, ny, nx), accumulated in place.
    Pinned by: identity and whole-voxel shifts reproduce the image, a 90
    degree turn is rot90, a random rigid motion equals scipy's order-1
    map_coordinates with zero padding, linearity in the frame weights
    (tests/test_oracle_motion.py)."""
    f = np.float32
    src = np.ascontiguousarray(img, f).reshape(-1)
    (nx, ny, nz), inv_v, origin, pos = _image_geometry(im)
    Z, Y, X = np.meshgrid(pos[2], pos[1], pos[0], indexing="ij")
    acc = np.ascontiguousarray(out, f).reshape(nz, ny, nx).copy()
    for r, fw in zip(np.asarray(inv_xf, f).reshape(-1, 12), np.asarray(weights, f)):
        q = [_fmaf(r[4 * a], X, _fmaf(r[4 * a + 1], Y, _fmaf(r[4 * a + 2], Z, r[4 * a + 3])))
             for a in range(3)]
        c = [((q[a] - origin[a]) * inv_v[a]).astype(f) for a in range(3)]
        i0 = [np.floor(c[a]).astype(np.int64) for a in range(3)]
        d = [(c[a] - i0[a].astype(f)).astype(f) for a in range(3)]
        one = f(1)
        for k in range(8):
            ox, oy, oz = k & 1, (k >> 1) & 1, (k >> 2) & 1
            wx = d[0] if ox else one - d[0]
            wy = d[1] if oy else one - d[1]
            wz = d[2] if oz else one - d[2]
            w = ((wx * wy) * wz).astype(f)
            x, y, z = i0[0] + ox, i0[1] + oy, i0[2] + oz
            ok = (x >= 0) & (x < nx) & (y >= 0) & (y < ny) & (z >= 0) & (z < nz)
            idx = np.where(ok, x + y * nx + z * nx * ny, 0)
            w = np.where(ok, w, f(0))
            acc = _fmaf((src[idx] * w).astype(f), fw, acc)
    return acc


def time_average_move(img, im: ImageDef, timestamps, xf, t0=None, t1=None):
    """util::timeAverageMoveImage(Device) (src/utils/ReconstructionUtils.cpp:
    497-592, ReconstructionUtilsDevice.cu:78-225): a zeroed output plus the
    weighted moved copies of frame 0 of `img` over [t0, t1] (default: the
    motion's full time range).  Returns (nz, ny, nx)."""
    if t0 is None:
        t0, t1 = full_time_range(timestamps)
    _, inv, w = time_average_frames(timestamps, xf, t0, t1)
    img = np.asarray(img, np.float32)
    base = img.reshape((-1,) + img.shape[-3:])[0]
    return move_accumulate(base, im, inv, w, np.zeros_like(base))


def time_average_move_dynamic(img, im: ImageDef, timestamps, xf, bounds):
    """util::timeAverageMoveImageDynamic (ReconstructionUtils.cpp:600-652):
    one time average per dynamic frame f over [bounds[f], bounds[f+1]].
    Returns (len(bounds) - 1, nz, ny, nx)."""
    b = [int(v) for v in bounds]
    return np.stack([time_average_move(img, im, timestamps, xf, b[f], b[f + 1])
                     for f in range(len(b) - 1)])


def sensitivity_images(sc: ScannerDef, im: ImageDef, proj: ProjDef,
                       num_subsets: int, listmode: bool, att_im=None,
                       att_img=None, hard_threshold=DEFAULT_HARD_THRESHOLD,
                       global_scale=1.0, dtype=np.float32, nthreads=1,
                       pos_ori=None, sens_histo=None, invert_sens=False,
                       image_psf=None, acf_histo=None):
    """OSEM::generateSensitivityImagesCore (src/recon/OSEM.cpp:926-986) with
    OSEM_CPU::generateSensitivityImageForCurrentSubset (OSEM_CPU.cpp:127-188):
    backproject the multiplicative correction (acf * sens, 1.0 without
    corrections; Corrector_CPU.cpp:17-45) of every bin of the scanner's
    uniform histogram (Corrector.cpp:152-157,228-241), one image per subset
    (one only for list-mode), projector without TOF (OSEM_CPU.cpp:94-108),
    then the hard threshold (voxels <= thr -> 0, others * global scale).
    sens_histo: per-bin normalisation (histogram order), inverted where
    nonzero when invert_sens (Corrector::getSensitivity, Corrector.cpp:
    295-308); the bin's factor is acf * sens (Corrector_CPU.cpp:17-45).
    image_psf (kx, ky, kz): the backprojection goes through the PSF's
    adjoint, then voxels whose unblurred value is <= EPS_FLT are zeroed
    (OSEM_CPU.cpp:158-165), before the hard threshold.  acf_histo: a
    hardware ACF histogram (histogram order), used per bin in place of the
    attenuation image (generateSensImage_kernel, OSEM_GPU.cu:1060-1090)."""
    pos, ori = pos_ori if pos_ori is not None else lut(sc)
    S = 1 if listmode else num_subsets
    p = proj.without_tof()
    im1 = ImageDef(im.nx, im.ny, im.nz, im.vx, im.vy, im.vz, im.ox, im.oy,
                   im.oz, 1, im.lx, im.ly, im.lz)
    out = []
    for s in range(S):
        bins = histo_subset_bins(sc, S, s)
        d1, d2 = histo_pairs(sc, bins)
        L = lors_from_pairs(pos, ori, d1, d2)
        if acf_histo is not None:
            vals = np.ascontiguousarray(np.asarray(acf_histo, np.float32).reshape(-1)[bins])
        elif att_img is not None:
            vals = attenuation_factors(att_im, att_img, L, nthreads)
        else:
            vals = np.ones(L.n, np.float32)
        if sens_histo is not None:
            sf = sensitivity_factors(sens_histo, bins, invert_sens)
            vals = (vals * sf).astype(np.float32)
        img = back(im1, p, L, vals, dtype=dtype, nthreads=nthreads)
        img = img.astype(np.float32) if dtype == np.float32 else img
        if image_psf is not None:
            blurred = psf_convolve(img, *image_psf, transpose=True)
            img = np.where(img <= EPS_FLT, blurred * 0, blurred).astype(np.float32)
        thr_img = img.copy()
        img = np.where(thr_img <= hard_threshold, img * 0, img * global_scale).astype(img.dtype)
        out.append(img)
    return out


def sensitivity_factors(sens_histo, bins, invert=False):
    """Corrector::getSensitivity (Corrector.cpp:295-308) for bins: the
    histogram value, 1/value where nonzero when inverted."""
    v = np.asarray(sens_histo, np.float32).reshape(-1)[np.asarray(bins, np.int64)]
    if invert:
        nz = v != 0
        v = v.copy()
        v[nz] = np.float32(1.0) / v[nz]
    return v.astype(np.float32)


def reconstruct(sc: ScannerDef, im: ImageDef, proj: ProjDef, lors: Lors,
                sens_images, num_iterations: int, num_subsets: int,
                listmode: bool, meas=None, att=None, rnd=None, scat=None,
                global_scale=1.0, subset_indices=None, dtype=np.float32,
                nthreads=1, initial=None, norm=None, image_psf=None, precorr=None):
    """OSEM::reconstruct / iterate (src/recon/OSEM.cpp:315-365,788-811) with
    the CPU hooks of OSEM_CPU.cpp:217-403:
      - list-mode sensitivity scaled by 1/S (OSEM.cpp:1007-1035);
      - image initialised to 0.125, masked where the (summed) sensitivity is
        <= 0 (OSEM_CPU.cpp:231-270);
      - per subset: u = A^T(m / (A x .* corr)) .* corr; x *= u/S where S > EPS.
    `lors` holds every event/bin; `subset_indices[s]` selects the subset
    (default: interleaved for list-mode, the caller supplies histogram
    subsets).  `att` is the per-LOR ACF (precomputeAttenuationFactors);
    `norm` the per-LOR normalisation factor (sensitivity_factors).
    image_psf (kx, ky, kz): the forward projection reads PSF(x) and the
    update image is PSF^T of the backprojected ratios
    (OSEM_CPU::computeEMUpdateImage, OSEM_CPU.cpp:284-295,377-392)."""
    dt = _dt(dtype)
    S = num_subsets
    sens = [np.asarray(s, dt).reshape(im.nz, im.ny, im.nx) for s in sens_images]
    if listmode and S != 1:
        sens = [sens[0] * dt(1.0 / S) if dt == np.float64 else
                (sens[0] * np.float32(np.float32(1.0) / np.float32(S))).astype(np.float32)]
    x = np.full(im.shape, INITIAL_VALUE_MLEM, dt) if initial is None else np.array(initial, dt)
    if listmode or S == 1:
        mask = sens[0]
    else:
        mask = np.zeros_like(sens[0])
        for s in sens:
            mask = mask + s
    x = np.where(mask[None] <= 0, x * 0, x).astype(dt)
    if subset_indices is None:
        subset_indices = [listmode_subset(lors.n, S, s) for s in range(S)]
    sub_lors = [lors.take(idx) for idx in subset_indices]
    take = lambda a, idx: None if a is None else np.ascontiguousarray(np.asarray(a)[idx], dt)
    for it in range(num_iterations):
        for s in range(S):
            idx = subset_indices[s]
            u = np.zeros(im.shape, dt)
            xf = x if image_psf is None else psf_convolve(x, *image_psf)
            em_subset(im, xf, u, proj, sub_lors[s], take(meas, idx), take(norm, idx),
                      take(att, idx), take(rnd, idx), take(scat, idx),
                      global_scale, DEFAULT_DENOM_THRESHOLD, dt, nthreads,
                      precorr=take(precorr, idx))
            if image_psf is not None:
                u = psf_convolve(u, *image_psf, transpose=True)
            sidx = 0 if listmode else s
            em_update(x, u, sens[sidx], dt(EPS_FLT))
    return x
