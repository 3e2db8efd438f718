"""Device-resident OSEM driver with the reference's OSEM interface
(include/yrt-pet/recon/OSEM.hpp:29-194; OSEM.cpp, OSEM_CPU.cpp, OSEM_GPU.cu).

Everything lives in HBM for the whole reconstruction: events / nonzero
histogram bins (SoA), per-LOR correction factors (computed once), the
sensitivity image(s), the image and the EM update image.  Per subset the
host issues: a fill of the update image, the EM kernels, one EM-update
kernel; there is no host-side property packing or per-batch copy.

With the Siddon projector (and `use_tile_plans`, the default) every subset
gets a tile plan once, before the first iteration: its rays' walks binned
by image tile, so that backprojection accumulates in LDS instead of issuing
one global atomic per voxel segment.

Data inputs:
  ListModeData   - d1, d2 (int32), optional tof (ps), dynamic frame,
                   motion frame + rigid transforms, randoms estimate.
  Histogram3DData - counts shaped (numZBin, numPhi, numR) in the
                   reference's Histogram3D layout.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import core

EPS_FLT = np.float32(1.19209e-07)                       # Constants.hpp:28


@dataclass
class ListModeData:
    """List-mode events (ListModeLUT: detector pair, TOF; dynamic framing and
    LOR motion as ListMode::addLORMotion / DynamicFraming)."""
    d1: torch.Tensor
    d2: torch.Tensor
    tof: torch.Tensor | None = None
    frame: torch.Tensor | None = None          # dynamic frame per event
    num_frames: int = 1
    motion_frame: torch.Tensor | None = None   # uint8 index into transforms
    motion_transforms: np.ndarray | None = None  # (k, 12) transform_t rows
    randoms: torch.Tensor | None = None

    def count(self):
        return self.d1.numel()

    def hasTOF(self):
        return self.tof is not None


@dataclass
class Histogram3DData:
    """Histogram3D counts (numZBin, numPhi, numR) (Histogram3D.cpp:165-203)."""
    values: torch.Tensor

    def count(self):
        return self.values.numel()


class OSEM:
    """OSEM with the reference's public members and setters.  GPU only."""

    DEFAULT_HARD_THRESHOLD = 1.0     # OSEM.hpp:34
    DEFAULT_DENOM_THRESHOLD = 1e-35  # OSEM.hpp:35
    INITIAL_VALUE_MLEM = 0.125       # OSEM.hpp:36

    def __init__(self, scanner: core.Scanner, device=0, ctx: core.Context | None = None):
        self.scanner = scanner
        self.ctx = ctx if ctx is not None else core.Context(scanner, device)
        self.dev = torch.device("cuda", self.ctx.device)
        self.num_MLEM_iterations = 10
        self.num_OSEM_subsets = 1
        self.hardThreshold = self.DEFAULT_HARD_THRESHOLD
        self.denomThreshold = self.DEFAULT_DENOM_THRESHOLD
        self.projectorParams = core.ProjectorParams(scanner)
        self.imageParams: core.ImageParams | None = None
        self.data = None
        self.globalScale = 1.0
        self.attImage = None           # mu (cm^-1), tensor
        self.attImageParams = None
        self.randomsHisto = None
        self.scatterHisto = None
        self.sensHisto = None
        self.invertSensitivity = False
        self.initialEstimate = None
        self.maskImage = None
        self.sensImages: list[torch.Tensor] = []
        self.lanes_per_event = None
        self.per_xcd = True
        self.use_tile_plans = True

    # ---- configuration (OSEM.hpp:57-134) ----
    def setImageParams(self, p):
        self.imageParams = p

    def getImageParams(self):
        return self.imageParams

    def setDataInput(self, d):
        self.data = d

    def isListModeEnabled(self):
        return isinstance(self.data, ListModeData)

    def setProjector(self, name):
        self.projectorParams.setProjector(name)

    def setNumRays(self, n):
        self.projectorParams.numRays = int(n)

    def addTOF(self, width_ps, n_std):
        self.projectorParams.addTOF(width_ps, n_std)

    def addProjPSF(self, fname_or_kernels, s_step=None, k_spacing=None):
        if isinstance(fname_or_kernels, str):
            self.projectorParams.setProjPsfFromFile(fname_or_kernels)
        else:
            self.projectorParams.psf_kernels = np.asarray(fname_or_kernels, np.float32)
            self.projectorParams.psf_s_step = float(s_step)
            self.projectorParams.psf_k_spacing = float(k_spacing)

    def setGlobalScalingFactor(self, g):
        self.globalScale = float(g)

    def setAttenuationImage(self, mu, params=None):
        """Attenuation image in cm^-1 (exp(-0.1 * integral), Tools.hpp:98-102)."""
        self.attImage = mu.contiguous()
        self.attImageParams = params or self.imageParams

    def setRandomsHistogram(self, h):
        self.randomsHisto = h

    def setScatterHistogram(self, h):
        self.scatterHisto = h

    def setSensitivityHistogram(self, h):
        self.sensHisto = h

    def setInvertSensitivity(self, invert=True):
        self.invertSensitivity = bool(invert)

    def setInitialEstimate(self, x):
        self.initialEstimate = x

    def setMaskImage(self, m):
        self.maskImage = m

    def setSensitivityImages(self, imgs):
        self.sensImages = [i.contiguous() for i in imgs]

    # ---- helpers ----
    def _projector(self, with_tof):
        p = self.ctx.projector(self.projectorParams, with_tof=with_tof)
        p.set_launch(self.lanes_per_event)
        p.set_accumulation(self.per_xcd)
        return p

    def _planned(self):
        return self.use_tile_plans and self.projectorParams.projectorType == "Siddon"

    def _sens_factor_per_bin(self, bins_flat):
        """Corrector::getSensitivity (Corrector.cpp:295-308)."""
        if self.sensHisto is None:
            return None
        s = self.sensHisto.reshape(-1)[bins_flat.long()].float()
        if self.invertSensitivity:
            s = torch.where(s != 0, 1.0 / s, s)
        return s.contiguous()

    def getExpectedSensImagesAmount(self):
        return 1 if self.isListModeEnabled() else self.num_OSEM_subsets

    # ---- sensitivity (OSEM.cpp:926-986, OSEM_CPU.cpp:127-188) ----
    def generateSensitivityImages(self):
        im = self.imageParams
        im1 = core.ImageParams(im.nx, im.ny, im.nz, im.vx, im.vy, im.vz, im.off_x,
                               im.off_y, im.off_z, 1, im.length_x, im.length_y, im.length_z)
        S = self.getExpectedSensImagesAmount()
        proj = self._projector(with_tof=False)
        out = []
        for s in range(S):
            img = torch.zeros(im1.shape, device=self.dev)
            L = core.LorSet.histo_subset(self.ctx, S, s)
            sens = None
            if self.sensHisto is not None:
                bins = self._subset_bins(S, s)
                sens = self._sens_factor_per_bin(bins)
                L = core.LorSet.bins_(bins)
            plan = self.ctx.plan(proj, im1, L) if self._planned() else None
            self.ctx.sensitivity(proj, im1, img, L, att_im=self.attImageParams,
                                 att_img=self.attImage, sens=sens, plan=plan)
            del plan
            # hard threshold: <= thr -> 0, else * global scale
            # (applyThreshold, OSEM_CPU.cpp:168-172)
            self.ctx.threshold(img, img.clone(), im1.nvox, self.hardThreshold, 0.0, 0.0,
                               self.globalScale, 0.0)
            out.append(img)
        self.sensImages = out
        return out

    def _subset_bins(self, S, s):
        """BinIteratorRangeHistogram3D(S, s) bin ids (BinIterator.cpp:191-198)
        as an int32 device tensor."""
        zb, ph, nr = self.ctx.histo_dims
        nps = ph // S
        if s + nps * S < ph:
            nps += 1
        idx = torch.arange(nr * nps * zb, device=self.dev, dtype=torch.int64)
        z = idx // (nps * nr)
        rem = idx % (nps * nr)
        phi = (rem // nr) * S + s
        r = rem % nr
        return (z * ph * nr + phi * nr + r).to(torch.int32)

    # ---- reconstruction (OSEM.cpp:315-365,788-811,1000-1035) ----
    def _prepare(self, proj, imt):
        d = self.data
        im = self.imageParams
        S = self.num_OSEM_subsets
        self.subsets = []
        if isinstance(d, ListModeData):
            if d.motion_transforms is not None:
                self.ctx.upload_motion(d.motion_transforms)
            n = d.count()
            att = None
            if self.attImage is not None:
                att = torch.empty(n, device=self.dev)
                Lall = core.LorSet.pairs(d.d1, d.d2, motion_frame=d.motion_frame)
                self.ctx.attenuation_factors(self.attImageParams, self.attImage, Lall, att)
            for s in range(S):
                # BinIteratorChronologicalInterleaved: events s, s+S, ...
                sl = slice(s, n, S)
                g = lambda t: None if t is None else t[sl].contiguous()
                L = core.LorSet.pairs(g(d.d1), g(d.d2), tof=g(d.tof), frame=g(d.frame),
                                      motion_frame=g(d.motion_frame))
                self.subsets.append(dict(L=L, meas=None, att=g(att), sens=None,
                                         randoms=g(d.randoms), scatter=None))
        else:
            vals = d.values.reshape(-1)
            for s in range(S):
                bins = self._subset_bins(S, s)
                m = vals[bins.long()].float()
                keep = m != 0  # zero-count bins contribute nothing (OSEM_GPU.cu:1224)
                bins = bins[keep].contiguous()
                m = m[keep].contiguous()
                L = core.LorSet.bins_(bins)
                att = None
                if self.attImage is not None:
                    att = torch.empty(bins.numel(), device=self.dev)
                    self.ctx.attenuation_factors(self.attImageParams, self.attImage, L, att)
                gb = lambda h: None if h is None else h.reshape(-1)[bins.long()].float().contiguous()
                self.subsets.append(dict(L=L, meas=m, att=att, sens=self._sens_factor_per_bin(bins),
                                         randoms=gb(self.randomsHisto), scatter=gb(self.scatterHisto)))
        for c in self.subsets:
            c["plan"] = self.ctx.plan(proj, imt, c["L"]) if self._planned() else None

    def reconstruct(self):
        im = self.imageParams
        d = self.data
        S = self.num_OSEM_subsets
        if len(self.sensImages) != self.getExpectedSensImagesAmount():
            raise RuntimeError(
                "The number of sensitivity images provided does not match the number "
                f"of subsets. Expected {self.getExpectedSensImagesAmount()} but "
                f"received {len(self.sensImages)}")
        # setupForDynamicRecon (OSEM.cpp:823-830)
        nt = d.num_frames if isinstance(d, ListModeData) else 1
        imt = core.ImageParams(im.nx, im.ny, im.nz, im.vx, im.vy, im.vz, im.off_x, im.off_y,
                               im.off_z, nt, im.length_x, im.length_y, im.length_z)
        sens = [s.reshape(-1) for s in self.sensImages]
        if self.isListModeEnabled() and S != 1:
            # initializeSensImageBuffer: x 1/S on a copy (OSEM.cpp:1007-1035)
            sens = [sens[0] * float(np.float32(1.0) / np.float32(S))]
        x = (self.initialEstimate.clone().reshape(imt.shape) if self.initialEstimate is not None
             else torch.full(imt.shape, self.INITIAL_VALUE_MLEM, device=self.dev))
        # mask (OSEM_CPU.cpp:241-269)
        if self.maskImage is not None:
            mask = self.maskImage.reshape(-1)
        elif S == 1 or self.isListModeEnabled():
            mask = sens[0]
        else:
            mask = torch.zeros_like(sens[0])
            for s in sens:
                mask += s
        self.ctx.threshold(x, mask.contiguous(), imt.nvox, 0.0, 0.0, 0.0, 1.0, 0.0)
        proj = self._projector(with_tof=True)
        self._prepare(proj, imt)
        u = torch.empty_like(x)
        for it in range(self.num_MLEM_iterations):
            for s in range(S):
                c = self.subsets[s]
                self.ctx.fill(u, 0.0)
                self.ctx.em_subset(proj, imt, x, u, c["L"], meas=c["meas"], sens=c["sens"],
                                   att=c["att"], randoms=c["randoms"], scatter=c["scatter"],
                                   global_scale=self.globalScale,
                                   denom_threshold=self.denomThreshold, plan=c["plan"])
                sidx = 0 if self.isListModeEnabled() else s
                self.ctx.em_update(x, u, sens[sidx].contiguous(), imt.nvox, EPS_FLT)
        return x


def createOSEM(scanner, useGPU=True, device=0):
    """util::createOSEM (src/utils/ReconstructionUtils.cpp:915-933); this
    build has the GPU implementation only."""
    if not useGPU:
        raise RuntimeError("yrtmi provides the GPU OSEM only")
    return OSEM(scanner, device)
