"""Shared test helpers: conversions between the product's geometry classes
and the oracle's, error metrics, and GPU tensor helpers."""
from __future__ import annotations

import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "yrt-pet_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import oracle as O  # noqa: E402


def oracle_scanner(sc):
    return O.ScannerDef(sc.axialFOV, sc.crystalSize_z, sc.crystalSize_trans,
                        sc.crystalDepth, sc.scannerRadius, sc.detsPerRing,
                        sc.numRings, sc.numDOI, sc.maxRingDiff, sc.minAngDiff,
                        sc.detsPerBlock)


def oracle_image(im):
    return O.ImageDef(im.nx, im.ny, im.nz, im.vx, im.vy, im.vz, im.off_x,
                      im.off_y, im.off_z, im.nt, im.length_x, im.length_y,
                      im.length_z)


def oracle_proj(pp, with_tof=True):
    sc = oracle_scanner(pp.scanner)
    return O.proj_for_scanner(
        sc, type="dd" if pp.projectorType == "DD" else "siddon",
        num_rays=pp.numRays,
        tof_width_ps=pp.tofWidth_ps if (with_tof and pp.hasTOF()) else None,
        tof_n_std=pp.tofNumStd, psf_kernels=pp.psf_kernels,
        psf_s_step=pp.psf_s_step, psf_k_spacing=pp.psf_k_spacing)


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
