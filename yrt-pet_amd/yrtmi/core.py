"""Thin Python layer over the yrtmi C ABI with torch tensors as device
buffers (torch is only used for device memory and streams).

Names mirror the reference's data model: ``Scanner`` (Scanner.cpp),
``ImageParams`` (ImageParams.cpp), ``ProjectorParams`` (ProjectorParams.hpp).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import capi

_P = C.c_void_p


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        if not t.is_cuda:
            raise ValueError("device tensor expected")
        if not t.is_contiguous():
            raise ValueError("contiguous tensor expected")
        return C.c_void_p(t.data_ptr())
    raise TypeError(type(t))


def _check_dtype(t, dt, name):
    if t is not None and t.dtype != dt:
        raise TypeError(f"{name} must be {dt}, got {t.dtype}")


# --------------------------------------------------------------------------
@dataclass
class Scanner:
    """Regular cylindrical scanner (Scanner ctor, src/datastruct/scanner/
    Scanner.cpp:101-127; the regular LUT of DetRegular.cpp:113-182)."""
    scannerName: str
    axialFOV: float
    crystalSize_z: float
    crystalSize_trans: float
    crystalDepth: float
    scannerRadius: float
    detsPerRing: int
    numRings: int
    numDOI: int = 1
    maxRingDiff: int = -1
    minAngDiff: int = 1
    detsPerBlock: int = 1
    _lut: tuple | None = field(default=None, repr=False)

    def __post_init__(self):
        if self.maxRingDiff < 0:
            self.maxRingDiff = self.numRings - 1
        if self.maxRingDiff >= self.numRings:
            raise ValueError("maxRingDiff must be < numRings")
        if self.minAngDiff <= 0:
            raise ValueError("minAngDiff must be > 0")

    def getNumDets(self):
        return self.detsPerRing * self.numRings * self.numDOI

    def lut(self):
        """(pos, orient) host arrays, n x 3 float32."""
        if self._lut is None:
            n = self.getNumDets()
            pos = np.zeros((n, 3), np.float32)
            ori = np.zeros((n, 3), np.float32)
            d = capi.ScannerDesc(self.axialFOV, self.crystalSize_z,
                                 self.crystalSize_trans, self.crystalDepth,
                                 self.scannerRadius, self.detsPerRing,
                                 self.numRings, self.numDOI, self.detsPerBlock)
            capi.check(capi.lib().yrtmi_lut_regular(
                C.byref(d), pos.ctypes.data_as(_P), ori.ctypes.data_as(_P)))
            self._lut = (pos, ori)
        return self._lut

    def histo_desc(self):
        return capi.HistoDesc(self.detsPerRing, self.numRings, self.numDOI,
                              self.maxRingDiff, self.minAngDiff)


@dataclass
class ImageParams:
    """Image geometry (ImageParams.cpp:136-155,206-266).  Build from voxel
    sizes (``vx``...) or from lengths with ``ImageParams.from_lengths``."""
    nx: int
    ny: int
    nz: int
    vx: float
    vy: float
    vz: float
    off_x: float = 0.0
    off_y: float = 0.0
    off_z: float = 0.0
    nt: int = 1
    length_x: float | None = None
    length_y: float | None = None
    length_z: float | None = None

    @staticmethod
    def from_lengths(nx, ny, nz, lx, ly, lz, ox=0.0, oy=0.0, oz=0.0, nt=1):
        f = np.float32
        return ImageParams(nx, ny, nz, float(f(lx) / f(nx)),
                           float(f(ly) / f(ny)), float(f(lz) / f(nz)), ox, oy,
                           oz, nt, float(f(lx)), float(f(ly)), float(f(lz)))

    def lengths(self):
        f = np.float32
        return (f(self.length_x) if self.length_x is not None else f(self.vx) * f(self.nx),
                f(self.length_y) if self.length_y is not None else f(self.vy) * f(self.ny),
                f(self.length_z) if self.length_z is not None else f(self.vz) * f(self.nz))

    @property
    def fovRadius(self):
        lx, ly, _ = self.lengths()
        r = max(lx / np.float32(2), ly / np.float32(2))
        return np.float32(r - max(np.float32(self.vx), np.float32(self.vy)) / np.float32(1000))

    @property
    def shape(self):
        return (self.nt, self.nz, self.ny, self.nx)

    @property
    def nvox(self):
        return self.nx * self.ny * self.nz

    def c(self, nt=None, zero_offset=False):
        lx, ly, lz = self.lengths()
        o = (0.0, 0.0, 0.0) if zero_offset else (self.off_x, self.off_y, self.off_z)
        return capi.Image(self.nx, self.ny, self.nz, self.nt if nt is None else nt,
                          self.vx, self.vy, self.vz, lx, ly, lz, o[0], o[1],
                          o[2], self.fovRadius)


@dataclass
class ProjectorParams:
    """ProjectorParams (include/yrt-pet/operators/ProjectorParams.hpp:27-71):
    projector type, numRays, TOF (width, n_std) and the projection-space PSF
    (one CSV row, ProjectionPsfManager.cpp:27-49, or explicit kernels)."""
    scanner: Scanner
    projectorType: str = "Siddon"
    numRays: int = 1
    tofWidth_ps: float | None = None
    tofNumStd: int = 0
    psf_kernels: np.ndarray | None = None
    psf_s_step: float = 0.0
    psf_k_spacing: float = 0.0
    # FP64 ray-plane setup (accuracy mode) instead of the reference FP32 order
    setupFP64: bool = False

    def setProjector(self, name):
        n = name.upper()
        if n not in ("S", "SIDDON", "D", "DD"):
            raise ValueError(f"unknown projector {name}")
        self.projectorType = "DD" if n.startswith("D") else "Siddon"

    def addTOF(self, tofWidth_ps, tofNumStd):
        self.tofWidth_ps = float(tofWidth_ps)
        self.tofNumStd = int(tofNumStd)

    def removeTOF(self):
        self.tofWidth_ps = None

    def hasTOF(self):
        return self.tofWidth_ps is not None

    def setProjPsfFromFile(self, fname):
        """One-row CSV: sStep, kSpacing, K, K*numS taps (ProjectionPsfManager::
        readFromFileInternal, ProjectionPsfManager.cpp:27-49)."""
        row = np.loadtxt(fname, delimiter=",", dtype=np.float32, ndmin=2)[0]
        K = int(row[2])
        if K % 2 == 0:
            raise RuntimeError("Kernels must be of odd length")
        num_s = (row.size - 3) // K
        self.psf_s_step = float(row[0])
        self.psf_k_spacing = float(row[1])
        self.psf_kernels = np.ascontiguousarray(row[3:3 + num_s * K].reshape(num_s, K))

    def desc(self, with_tof=True):
        ker = None if self.psf_kernels is None else np.ascontiguousarray(self.psf_kernels, np.float32)
        self._keep = ker
        return capi.ProjectorDesc(
            capi.PROJ_DD if self.projectorType == "DD" else capi.PROJ_SIDDON,
            int(self.numRays), 1 if (with_tof and self.hasTOF()) else 0,
            float(self.tofWidth_ps or 0.0), int(self.tofNumStd),
            None if ker is None else ker.ctypes.data_as(_P),
            0 if ker is None else ker.shape[0], 0 if ker is None else ker.shape[1],
            self.psf_s_step, self.psf_k_spacing, self.scanner.crystalSize_z,
            self.scanner.crystalSize_trans, self.scanner.crystalDepth,
            1 if self.setupFP64 else 0)


# --------------------------------------------------------------------------
class LorSet:
    """A device-resident set of LORs (SoA tensors) plus how to resolve them:
    detector pairs (list-mode), histogram bin ids, or a histogram subset
    iterator."""

    def __init__(self, kind, n, d1=None, d2=None, bins=None, num_subsets=1,
                 subset=0, tof=None, frame=None, motion_frame=None):
        _check_dtype(d1, torch.int32, "d1")
        _check_dtype(d2, torch.int32, "d2")
        _check_dtype(bins, torch.int32, "bins")
        _check_dtype(tof, torch.float32, "tof")
        _check_dtype(frame, torch.int32, "frame")
        _check_dtype(motion_frame, torch.uint8, "motion_frame")
        for t in (d1, d2, bins, tof, frame, motion_frame):
            if t is not None and t.numel() < n:
                raise ValueError("LOR arrays shorter than n")
        self.kind, self.n = kind, int(n)
        self.d1, self.d2, self.bins = d1, d2, bins
        self.num_subsets, self.subset = num_subsets, subset
        self.tof, self.frame, self.motion_frame = tof, frame, motion_frame

    @staticmethod
    def pairs(d1, d2, tof=None, frame=None, motion_frame=None):
        return LorSet(capi.SRC_PAIRS, d1.numel(), d1=d1, d2=d2, tof=tof,
                      frame=frame, motion_frame=motion_frame)

    @staticmethod
    def bins_(bins, tof=None, frame=None):
        return LorSet(capi.SRC_BINS, bins.numel(), bins=bins, tof=tof, frame=frame)

    @staticmethod
    def histo_subset(ctx, num_subsets, subset):
        n = capi.lib().yrtmi_histo_subset_size(ctx.h, num_subsets, subset)
        if n < 0:
            raise capi.YrtmiError("histogram not set up or bad subset")
        return LorSet(capi.SRC_HISTO_SUBSET, n, num_subsets=num_subsets, subset=subset)

    def slice(self, start, stop):
        """Contiguous sub-range (pairs / bins sources)."""
        if self.kind == capi.SRC_HISTO_SUBSET:
            raise ValueError("cannot slice a subset iterator")
        g = (lambda t: None if t is None else t[start:stop])
        return LorSet(self.kind, stop - start, g(self.d1), g(self.d2), g(self.bins),
                      self.num_subsets, self.subset, g(self.tof), g(self.frame),
                      g(self.motion_frame))

    def permuted(self, ctx, perm):
        """The same LORs in the order perm (pairs / bins sources): element i
        of every per-LOR array is element perm[i] of this set's."""
        if self.kind == capi.SRC_HISTO_SUBSET:
            raise ValueError("cannot permute a subset iterator")
        if perm.numel() != self.n:
            raise ValueError("permutation size differs from the LOR count")
        g = (lambda t: None if t is None else ctx.permute(t[:self.n], perm))
        return LorSet(self.kind, self.n, g(self.d1), g(self.d2), g(self.bins),
                      self.num_subsets, self.subset, g(self.tof), g(self.frame),
                      g(self.motion_frame))

    def c(self):
        return capi.Lors(self.kind, self.n, _ptr(self.d1), _ptr(self.d2),
                         _ptr(self.bins), self.num_subsets, self.subset,
                         _ptr(self.tof), _ptr(self.frame), _ptr(self.motion_frame))


class Projector:
    def __init__(self, ctx, params: ProjectorParams, with_tof=True):
        self.ctx = ctx
        self.params = params
        h = _P()
        d = params.desc(with_tof)
        capi.check(capi.lib().yrtmi_proj_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self.num_rays = params.numRays if params.projectorType == "Siddon" else 1

    def set_launch(self, lanes_per_event=None, block=256, grid=0):
        if lanes_per_event is None:
            r = self.num_rays
            lanes_per_event = 4 if r >= 4 else (2 if r >= 2 else 1)
        capi.check(capi.lib().yrtmi_proj_set_launch(self.h, lanes_per_event, block, grid))

    def set_accumulation(self, per_xcd=True):
        """Backprojection accumulation: per-XCD private images folded after
        the kernel (default) or direct atomics into the target."""
        capi.check(capi.lib().yrtmi_proj_set_accumulation(self.h, 1 if per_xcd else 0))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                capi.lib().yrtmi_proj_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Plan:
    """Tile plan of a LOR set (yrtmi_plan_create): the Siddon walks of its
    rays (or its DD footprints) binned by image tiles of `tile_shape`, and
    its ray table.  Geometry only; holds device memory until dropped.  Keeps
    the LOR set alive (the plan refers to its arrays)."""

    def __init__(self, ctx, proj, im, lors, chunk_records=0, tile_shape=0, persistent=0,
                 single_use=False):
        self.ctx, self.proj, self.im, self.lors = ctx, proj, im, lors
        h = _P()
        ci, cl = im.c(), lors.c()
        opt = capi.PlanOptions(int(chunk_records), int(tile_shape), int(persistent), 0,
                               int(bool(single_use)))
        capi.check(capi.lib().yrtmi_plan_create(ctx.h, proj.h, C.byref(ci), C.byref(cl),
                                                C.byref(opt), C.byref(h)))
        self.h = h
        self.tile_shape = tile_shape

    def set_persistent(self, workgroups):
        """0: one workgroup per chunk; n > 0: n workgroups on the chunk queue."""
        capi.check(capi.lib().yrtmi_plan_set_persistent(self.h, int(workgroups)))

    def info(self):
        r, c, b = C.c_int64(), C.c_int64(), C.c_int64()
        capi.check(capi.lib().yrtmi_plan_info(self.h, C.byref(r), C.byref(c), C.byref(b)))
        return dict(records=r.value, chunks=c.value, bytes=b.value)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                capi.lib().yrtmi_plan_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Graph:
    """The calls made on a context inside ``with ctx.graph() as g:`` are
    recorded (not run) and replayed by ``g.launch()`` (yrtmi_graph_*)."""

    def __init__(self, ctx):
        self.ctx, self.h = ctx, None

    def __enter__(self):
        capi.check(capi.lib().yrtmi_graph_begin(self.ctx.h))
        return self

    def __exit__(self, et, ev, tb):
        h = _P()
        rc = capi.lib().yrtmi_graph_end(self.ctx.h, C.byref(h))
        if rc == 0:
            self.h = h
        if et is None:
            capi.check(rc)
        return False

    def launch(self):
        capi.check(capi.lib().yrtmi_graph_launch(self.ctx.h, self.h))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                capi.lib().yrtmi_graph_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Context:
    """A yrtmi context bound to one GPU and a HIP stream (torch's current
    stream by default), holding the resident detector LUT and histogram /
    motion tables of one scanner."""

    def __init__(self, scanner: Scanner, device=0, stream=None, histogram=True):
        self.device = device
        torch.cuda.set_device(device)
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self.stream = stream
        h = _P()
        capi.check(capi.lib().yrtmi_ctx_create(device, _P(stream.cuda_stream), C.byref(h)))
        self.h = h
        self.scanner = scanner
        pos, ori = scanner.lut()
        capi.check(capi.lib().yrtmi_lut_upload(h, pos.ctypes.data_as(_P),
                                               ori.ctypes.data_as(_P), pos.shape[0]))
        self.histo_dims = None
        if histogram and scanner.detsPerRing % 2 == 0:
            z, p, r = C.c_int64(), C.c_int64(), C.c_int64()
            d = scanner.histo_desc()
            capi.check(capi.lib().yrtmi_histo_setup(h, C.byref(d), C.byref(z),
                                                    C.byref(p), C.byref(r)))
            self.histo_dims = (z.value, p.value, r.value)

    def set_stream(self, stream):
        self.stream = stream
        capi.check(capi.lib().yrtmi_ctx_set_stream(self.h, _P(stream.cuda_stream)))

    def upload_motion(self, xf):
        xf = np.ascontiguousarray(xf, np.float32).reshape(-1, 12)
        capi.check(capi.lib().yrtmi_motion_upload(self.h, xf.ctypes.data_as(_P), xf.shape[0]))

    def synchronize(self):
        capi.check(capi.lib().yrtmi_ctx_synchronize(self.h))

    def projector(self, params, with_tof=True):
        return Projector(self, params, with_tof)

    # ---- entry points ----
    def forward(self, proj: Projector, im: ImageParams, img, lors: LorSet, out):
        _check_dtype(img, torch.float32, "img")
        _check_dtype(out, torch.float32, "out")
        if img.numel() != im.nvox * im.nt or out.numel() < lors.n:
            raise ValueError("forward: size mismatch")
        ci, cl = im.c(), lors.c()
        capi.check(capi.lib().yrtmi_forward(self.h, proj.h, C.byref(ci), _ptr(img),
                                            C.byref(cl), _ptr(out)))
        return out

    def graph(self):
        """Record the calls of a ``with`` block as one replayable graph; the
        context's stream must not be the null stream (torch's default
        stream is: use ``set_stream(torch.cuda.Stream())``)."""
        return Graph(self)

    def autotune(self, proj: Projector, im: ImageParams, lors: LorSet, x):
        """Time launch shapes on ``lors`` / image ``x`` (yrtmi_autotune): sets
        the projector's lanes per event, returns the fastest settings."""
        _check_dtype(x, torch.float32, "x")
        t = capi.Tuning()
        ci, cl = im.c(), lors.c()
        capi.check(capi.lib().yrtmi_autotune(self.h, proj.h, C.byref(ci), C.byref(cl), _ptr(x),
                                             C.byref(t)))
        return dict(lanes_per_event=t.lanes_per_event, tile_shape=t.tile_shape,
                    persistent=t.persistent, ratio_ms=t.ratio_ms, back_ms=t.back_ms,
                    ratio_ms_lpe=list(t.ratio_ms_lpe),
                    back_ms_shape=[list(r) for r in t.back_ms_shape])

    def plan(self, proj: Projector, im: ImageParams, lors: LorSet, chunk_records=0,
             tile_shape=0, persistent=0, single_use=False):
        """single_use: a plan backprojected once (e.g. for a sensitivity
        image), built without ordering its records by length."""
        return Plan(self, proj, im, lors, chunk_records, tile_shape, persistent, single_use)

    def back(self, proj: Projector, im: ImageParams, img, lors: LorSet, vals, plan=None):
        _check_dtype(img, torch.float32, "img")
        _check_dtype(vals, torch.float32, "vals")
        if img.numel() != im.nvox * im.nt or vals.numel() < lors.n:
            raise ValueError("back: size mismatch")
        ci, cl = im.c(), lors.c()
        if plan is not None:
            capi.check(capi.lib().yrtmi_back_planned(self.h, proj.h, C.byref(ci), _ptr(img),
                                                     C.byref(cl), _ptr(vals), plan.h))
        else:
            capi.check(capi.lib().yrtmi_back(self.h, proj.h, C.byref(ci), _ptr(img),
                                             C.byref(cl), _ptr(vals)))
        return img

    def em_subset(self, proj, im, x, u, lors, meas=None, sens=None, att=None,
                  randoms=None, scatter=None, global_scale=1.0,
                  denom_threshold=1e-35, plan=None):
        for t, nm in ((x, "x"), (u, "u"), (meas, "meas"), (sens, "sens"),
                      (att, "att"), (randoms, "randoms"), (scatter, "scatter")):
            _check_dtype(t, torch.float32, nm)
        if x.numel() != im.nvox * im.nt or u.numel() != x.numel():
            raise ValueError("em_subset: size mismatch")
        for t in (meas, sens, att, randoms, scatter):
            if t is not None and t.numel() < lors.n:
                raise ValueError("em_subset: correction array shorter than the LOR set")
        corr = capi.Corrections(_ptr(meas), _ptr(sens), _ptr(att), _ptr(randoms),
                                _ptr(scatter), global_scale, denom_threshold)
        ci, cl = im.c(), lors.c()
        if plan is not None:
            capi.check(capi.lib().yrtmi_em_subset_planned(
                self.h, proj.h, C.byref(ci), _ptr(x), _ptr(u), C.byref(cl), C.byref(corr),
                plan.h))
        else:
            capi.check(capi.lib().yrtmi_em_subset(self.h, proj.h, C.byref(ci), _ptr(x),
                                                  _ptr(u), C.byref(cl), C.byref(corr)))
        return u

    def em_ratio(self, proj, im, x, lors, v, meas=None, sens=None, att=None,
                 randoms=None, scatter=None, global_scale=1.0, denom_threshold=1e-35,
                 plan=None):
        """v[i] = EM ratio of LOR i (the forward half of em_subset); with a
        plan of (proj, im, lors) the LORs' rays come from its ray table."""
        for t, nm in ((x, "x"), (v, "v"), (meas, "meas"), (sens, "sens"),
                      (att, "att"), (randoms, "randoms"), (scatter, "scatter")):
            _check_dtype(t, torch.float32, nm)
        if x.numel() != im.nvox * im.nt or v.numel() < lors.n:
            raise ValueError("em_ratio: size mismatch")
        for t in (meas, sens, att, randoms, scatter):
            if t is not None and t.numel() < lors.n:
                raise ValueError("em_ratio: correction array shorter than the LOR set")
        corr = capi.Corrections(_ptr(meas), _ptr(sens), _ptr(att), _ptr(randoms),
                                _ptr(scatter), global_scale, denom_threshold)
        ci, cl = im.c(), lors.c()
        if plan is not None:
            capi.check(capi.lib().yrtmi_em_ratio_planned(self.h, proj.h, C.byref(ci), _ptr(x),
                                                         C.byref(cl), C.byref(corr), _ptr(v),
                                                         plan.h))
        else:
            capi.check(capi.lib().yrtmi_em_ratio(self.h, proj.h, C.byref(ci), _ptr(x),
                                                 C.byref(cl), C.byref(corr), _ptr(v)))
        return v

    def sensitivity(self, proj, im, img, lors, att_im=None, att_img=None, sens=None,
                    plan=None):
        _check_dtype(img, torch.float32, "img")
        if img.numel() != im.nvox:
            raise ValueError("sensitivity: image must be single-frame")
        ci, cl = im.c(nt=1), lors.c()
        ca = att_im.c(nt=1) if att_img is not None else None
        if att_img is not None and att_img.numel() != att_im.nvox:
            raise ValueError("sensitivity: attenuation image size mismatch")
        if plan is not None:
            capi.check(capi.lib().yrtmi_sensitivity_planned(
                self.h, proj.h, C.byref(ci), _ptr(img), C.byref(cl),
                C.byref(ca) if ca is not None else None, _ptr(att_img), _ptr(sens), plan.h))
        else:
            capi.check(capi.lib().yrtmi_sensitivity(
                self.h, proj.h, C.byref(ci), _ptr(img), C.byref(cl),
                C.byref(ca) if ca is not None else None, _ptr(att_img), _ptr(sens)))
        return img

    def histo_subset_nonzero(self, values, num_subsets, subset):
        """The nonzero bins of histogram subset `subset` of `num_subsets`
        (BinIteratorRangeHistogram3D order) and their values: (int32 bin ids,
        float32 values), each of the nonzero count (yrtmi_histo_subset_nonzero;
        the reference GPU EM skips zero-measurement bins)."""
        _check_dtype(values, torch.float32, "values")
        cap = capi.lib().yrtmi_histo_subset_size(self.h, num_subsets, subset)
        if cap < 0:
            raise capi.YrtmiError("histogram not set up or bad subset")
        dev = values.device
        bins = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        meas = torch.empty(max(cap, 1), dtype=torch.float32, device=dev)
        n = C.c_int64()
        capi.check(capi.lib().yrtmi_histo_subset_nonzero(self.h, _ptr(values), num_subsets,
                                                         subset, _ptr(bins), _ptr(meas),
                                                         C.byref(n)))
        return bins[:n.value].clone(), meas[:n.value].clone()

    def histo_pair_bins(self, d1, d2):
        """The Histogram3D bin of each detector pair (int32 tensor of bin
        ids; yrtmi_histo_pair_bins, Histogram3D::getBinIdFromDetPair)."""
        _check_dtype(d1, torch.int32, "d1")
        _check_dtype(d2, torch.int32, "d2")
        if d1.numel() != d2.numel():
            raise ValueError("histo_pair_bins: d1 and d2 differ in length")
        out = torch.empty(d1.numel(), dtype=torch.int32, device=d1.device)
        capi.check(capi.lib().yrtmi_histo_pair_bins(self.h, _ptr(d1), _ptr(d2), d1.numel(),
                                                    _ptr(out)))
        return out

    def attenuation_factors(self, att_im, att_img, lors, out):
        if att_img.numel() != att_im.nvox or out.numel() < lors.n:
            raise ValueError("attenuation: size mismatch")
        ca, cl = att_im.c(nt=1), lors.c()
        capi.check(capi.lib().yrtmi_attenuation_factors(self.h, C.byref(ca), _ptr(att_img),
                                                        C.byref(cl), _ptr(out)))
        return out

    def em_update(self, x, u, s, nvox, thr=np.float32(1.19209e-07)):
        nt = x.numel() // nvox
        sf = s.numel() // nvox
        if x.numel() != nt * nvox or u.numel() != x.numel() or s.numel() != sf * nvox:
            raise ValueError("em_update: size mismatch")
        capi.check(capi.lib().yrtmi_em_update(self.h, _ptr(x), _ptr(u), _ptr(s),
                                              nvox, nt, sf, float(thr)))
        return x

    def threshold(self, x, mask, nvox, thr, le_scale, le_off, gt_scale, gt_off):
        nt = x.numel() // nvox
        mf = mask.numel() // nvox
        capi.check(capi.lib().yrtmi_threshold(self.h, _ptr(x), _ptr(mask), nvox, nt, mf,
                                              thr, le_scale, le_off, gt_scale, gt_off))
        return x

    def fill(self, x, v):
        capi.check(capi.lib().yrtmi_fill(self.h, _ptr(x), x.numel(), float(v)))
        return x

    def psf(self, im, x, out, tmp, kx, ky, kz, transpose=False):
        """out = image-space PSF of x (separable circular convolution with
        kx, ky, kz; the adjoint with transpose).  tmp: a scratch image."""
        for t, nm in ((x, "x"), (out, "out"), (tmp, "tmp")):
            _check_dtype(t, torch.float32, nm)
            if t.numel() != im.nvox * im.nt:
                raise ValueError(f"psf: {nm} size mismatch")
        ks = [np.ascontiguousarray(k, np.float32) for k in (kx, ky, kz)]
        ci = im.c()
        f = C.POINTER(C.c_float)
        capi.check(capi.lib().yrtmi_psf_apply(
            self.h, C.byref(ci), _ptr(x), _ptr(out), _ptr(tmp),
            ks[0].ctypes.data_as(f), len(ks[0]), ks[1].ctypes.data_as(f), len(ks[1]),
            ks[2].ctypes.data_as(f), len(ks[2]), 1 if transpose else 0))
        return out

    def time_average_move(self, im, x, out, transforms, weights):
        """out (one frame of im) += sum over frames f of weights[f] * x
        (its frame 0) moved by transforms[f] (k x 12 transform_t rows; the
        library inverts them) with trilinear interpolation."""
        _check_dtype(x, torch.float32, "x")
        _check_dtype(out, torch.float32, "out")
        if x.numel() < im.nvox or out.numel() < im.nvox:
            raise ValueError("time_average_move: x and out need one frame each")
        xf = np.ascontiguousarray(transforms, np.float32).reshape(-1, 12)
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        if len(w) != len(xf):
            raise ValueError("time_average_move: one weight per transform")
        ci = im.c()
        f = C.POINTER(C.c_float)
        capi.check(capi.lib().yrtmi_time_average_move(
            self.h, C.byref(ci), _ptr(x), _ptr(out), xf.ctypes.data_as(f),
            w.ctypes.data_as(f), len(w)))
        return out

    def lor_order(self, im, lors):
        """Locality order of `lors` (int32 tensor of the indices 0..n-1):
        sorted by frame, axial midpoint, tilt, transaxial angle and offset,
        so that LORs projected together read the same voxels."""
        perm = torch.empty(lors.n, dtype=torch.int32, device=torch.device("cuda", self.device))
        ci, cl = im.c(), lors.c()
        capi.check(capi.lib().yrtmi_lor_order(self.h, C.byref(ci), C.byref(cl), _ptr(perm)))
        return perm

    def permute(self, src, perm):
        """src[perm] (a new contiguous tensor) on the context's stream."""
        if src.dim() != 1 or not src.is_contiguous() or src.numel() < perm.numel():
            raise ValueError("permute: src must be a contiguous vector of at least n elements")
        _check_dtype(perm, torch.int32, "perm")
        out = torch.empty(perm.numel(), dtype=src.dtype, device=src.device)
        capi.check(capi.lib().yrtmi_permute(self.h, _ptr(src), src.element_size(), _ptr(perm),
                                            perm.numel(), _ptr(out)))
        return out

    def count_segments(self, proj, im, lors):
        tot = C.c_int64()
        ci, cl = im.c(), lors.c()
        capi.check(capi.lib().yrtmi_count_segments(self.h, proj.h, C.byref(ci),
                                                   C.byref(cl), C.byref(tot)))
        return tot.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                capi.lib().yrtmi_ctx_destroy(self.h)
                self.h = None
        except Exception:
            pass
