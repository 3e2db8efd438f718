"""yrtmi - MI355X-native OSEM projection path (Siddon, multi-ray Siddon and
distance-driven projectors with TOF and projection-space PSF, sensitivity
generation and the OSEM subset loop) behind the reference's operator / OSEM
names.  Hand-written HIP kernels for gfx950 behind a C ABI (include/yrtmi.h).
"""
from . import capi
from .core import Context, Graph, ImageParams, LorSet, Plan, Projector, ProjectorParams, Scanner

__all__ = ["capi", "Context", "Graph", "ImageParams", "LorSet", "Plan", "Projector",
           "ProjectorParams", "Scanner"]
