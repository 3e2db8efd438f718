"""ctypes binding of the yrtmi C ABI (include/yrtmi.h).

This module loads ``libyrtmi.so`` (built in-tree next to this file) and
raises ``RuntimeError`` when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libyrtmi.so")

PROJ_SIDDON = 0
PROJ_DD = 1
SRC_PAIRS = 0
SRC_BINS = 1
SRC_HISTO_SUBSET = 2
MAX_RAYS = 16


class YrtmiError(RuntimeError):
    pass


class Image(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("nt", C.c_int64), ("vx", C.c_float), ("vy", C.c_float),
                ("vz", C.c_float), ("lx", C.c_float), ("ly", C.c_float),
                ("lz", C.c_float), ("ox", C.c_float), ("oy", C.c_float),
                ("oz", C.c_float), ("fov_radius", C.c_float)]


class ProjectorDesc(C.Structure):
    _fields_ = [("type", C.c_int), ("num_rays", C.c_int),
                ("use_tof", C.c_int), ("tof_width_ps", C.c_float),
                ("tof_n_std", C.c_int), ("psf_kernels", C.c_void_p),
                ("psf_num_s", C.c_int), ("psf_ksize", C.c_int),
                ("psf_s_step", C.c_float), ("psf_k_spacing", C.c_float),
                ("crystal_size_z", C.c_float),
                ("crystal_size_trans", C.c_float),
                ("crystal_depth", C.c_float), ("setup_fp64", C.c_int)]


class HistoDesc(C.Structure):
    _fields_ = [("dets_per_ring", C.c_int64), ("num_rings", C.c_int64),
                ("num_doi", C.c_int64), ("max_ring_diff", C.c_int64),
                ("min_ang_diff", C.c_int64)]


class Lors(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int64), ("d1", C.c_void_p),
                ("d2", C.c_void_p), ("bins", C.c_void_p),
                ("num_subsets", C.c_int), ("subset", C.c_int),
                ("tof", C.c_void_p), ("frame", C.c_void_p),
                ("motion_frame", C.c_void_p)]


class Corrections(C.Structure):
    _fields_ = [("meas", C.c_void_p), ("sens", C.c_void_p),
                ("att", C.c_void_p), ("randoms", C.c_void_p),
                ("scatter", C.c_void_p), ("global_scale", C.c_float),
                ("denom_threshold", C.c_float), ("precorr", C.c_void_p)]


class ScannerDesc(C.Structure):
    _fields_ = [("axial_fov", C.c_float), ("crystal_size_z", C.c_float),
                ("crystal_size_trans", C.c_float),
                ("crystal_depth", C.c_float), ("radius", C.c_float),
                ("dets_per_ring", C.c_int64), ("num_rings", C.c_int64),
                ("num_doi", C.c_int64), ("dets_per_block", C.c_int64)]


_lib = None

_P = C.c_void_p
class PlanOptions(C.Structure):
    _fields_ = [("chunk_records", C.c_int), ("tile_shape", C.c_int), ("persistent", C.c_int),
                ("max_bytes", C.c_int64), ("single_use", C.c_int)]


TILE_16x16x16, TILE_32x32x16, TILE_32x32x32, TILE_32x32x8 = 0, 1, 2, 3


class Tuning(C.Structure):
    _fields_ = [("lanes_per_event", C.c_int), ("tile_shape", C.c_int), ("persistent", C.c_int),
                ("ratio_ms", C.c_float), ("back_ms", C.c_float),
                ("ratio_ms_lpe", C.c_float * 4), ("back_ms_shape", (C.c_float * 2) * 4)]

_SIGS = {
    "yrtmi_last_error": (C.c_char_p, []),
    "yrtmi_ctx_create": (C.c_int, [C.c_int, _P, C.POINTER(_P)]),
    "yrtmi_ctx_destroy": (C.c_int, [_P]),
    "yrtmi_ctx_set_stream": (C.c_int, [_P, _P]),
    "yrtmi_ctx_synchronize": (C.c_int, [_P]),
    "yrtmi_lut_upload": (C.c_int, [_P, _P, _P, C.c_int64]),
    "yrtmi_lut_regular": (C.c_int, [C.POINTER(ScannerDesc), _P, _P]),
    "yrtmi_histo_setup": (C.c_int, [_P, C.POINTER(HistoDesc), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "yrtmi_histo_subset_size": (C.c_int64, [_P, C.c_int, C.c_int]),
    "yrtmi_motion_upload": (C.c_int, [_P, _P, C.c_int]),
    "yrtmi_proj_create": (C.c_int, [_P, C.POINTER(ProjectorDesc), C.POINTER(_P)]),
    "yrtmi_proj_destroy": (C.c_int, [_P]),
    "yrtmi_proj_set_launch": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "yrtmi_proj_set_accumulation": (C.c_int, [_P, C.c_int]),
    "yrtmi_forward": (C.c_int, [_P, _P, C.POINTER(Image), _P, C.POINTER(Lors), _P]),
    "yrtmi_back": (C.c_int, [_P, _P, C.POINTER(Image), _P, C.POINTER(Lors), _P]),
    "yrtmi_em_subset": (C.c_int, [_P, _P, C.POINTER(Image), _P, _P,
                                  C.POINTER(Lors), C.POINTER(Corrections)]),
    "yrtmi_em_subset_noalloc": (C.c_int, [_P, _P, C.POINTER(Image), _P, _P,
                                          C.POINTER(Lors), C.POINTER(Corrections)]),
    "yrtmi_sensitivity": (C.c_int, [_P, _P, C.POINTER(Image), _P, C.POINTER(Lors),
                                    C.POINTER(Image), _P, _P]),
    "yrtmi_attenuation_factors": (C.c_int, [_P, C.POINTER(Image), _P,
                                            C.POINTER(Lors), _P]),
    "yrtmi_em_update": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_int64,
                                  C.c_int64, C.c_float]),
    "yrtmi_threshold": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int64, C.c_int64,
                                  C.c_float, C.c_float, C.c_float, C.c_float,
                                  C.c_float]),
    "yrtmi_fill": (C.c_int, [_P, _P, C.c_int64, C.c_float]),
    "yrtmi_psf_apply": (C.c_int, [_P, C.POINTER(Image), _P, _P, _P, _P, C.c_int, _P, C.c_int,
                                  _P, C.c_int, C.c_int]),
    "yrtmi_histo_gather_pairs": (C.c_int, [_P, _P, _P, _P, C.c_int64, _P]),
    "yrtmi_histo_pair_bins": (C.c_int, [_P, _P, _P, C.c_int64, _P]),
    "yrtmi_fill_i32": (C.c_int, [_P, _P, C.c_int64, C.c_int32]),
    "yrtmi_mem_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "yrtmi_time_average_move": (C.c_int, [_P, C.POINTER(Image), _P, _P, _P, _P, C.c_int]),
    "yrtmi_count_segments": (C.c_int, [_P, _P, C.POINTER(Image), C.POINTER(Lors),
                                       C.POINTER(C.c_int64)]),
    "yrtmi_forward_f64": (C.c_int, [_P, _P, C.POINTER(Image), _P, C.POINTER(Lors), _P]),
    "yrtmi_back_f64": (C.c_int, [_P, _P, C.POINTER(Image), _P, C.POINTER(Lors), _P]),
    "yrtmi_device_alloc": (C.c_int, [_P, C.c_int64, C.POINTER(_P)]),
    "yrtmi_device_free": (C.c_int, [_P, _P]),
    "yrtmi_copy": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int]),
    "yrtmi_histo_subset_nonzero": (C.c_int, [_P, _P, C.c_int, C.c_int, _P, _P,
                                             C.POINTER(C.c_int64)]),
    "yrtmi_subset_interleaved": (C.c_int, [_P, _P, C.c_int, C.c_int64, C.c_int, C.c_int, _P,
                                           C.POINTER(C.c_int64)]),
    "yrtmi_gather": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int64, C.c_int, _P]),
    "yrtmi_histo_subset_gather": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, _P]),
    "yrtmi_axpby": (C.c_int, [_P, C.c_int64, C.c_float, _P, C.c_float, _P]),
    "yrtmi_lor_order": (C.c_int, [_P, C.POINTER(Image), C.POINTER(Lors), _P]),
    "yrtmi_permute": (C.c_int, [_P, _P, C.c_int, _P, C.c_int64, _P]),
    "yrtmi_plan_create": (C.c_int, [_P, _P, C.POINTER(Image), C.POINTER(Lors),
                                    C.POINTER(PlanOptions), C.POINTER(_P)]),
    "yrtmi_plan_set_persistent": (C.c_int, [_P, C.c_int]),
    "yrtmi_ctx_own_stream": (C.c_int, [_P, C.POINTER(_P)]),
    "yrtmi_autotune": (C.c_int, [_P, _P, C.POINTER(Image), C.POINTER(Lors), _P,
                                 C.POINTER(Tuning)]),
    "yrtmi_graph_begin": (C.c_int, [_P]),
    "yrtmi_graph_end": (C.c_int, [_P, C.POINTER(_P)]),
    "yrtmi_graph_launch": (C.c_int, [_P, _P]),
    "yrtmi_graph_destroy": (C.c_int, [_P]),
    "yrtmi_plan_destroy": (C.c_int, [_P]),
    "yrtmi_plan_info": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64)]),
    "yrtmi_back_planned": (C.c_int, [_P, _P, C.POINTER(Image), _P, C.POINTER(Lors), _P, _P]),
    "yrtmi_em_subset_planned": (C.c_int, [_P, _P, C.POINTER(Image), _P, _P,
                                          C.POINTER(Lors), C.POINTER(Corrections), _P]),
    "yrtmi_em_ratio": (C.c_int, [_P, _P, C.POINTER(Image), _P, C.POINTER(Lors),
                                 C.POINTER(Corrections), _P]),
    "yrtmi_em_ratio_planned": (C.c_int, [_P, _P, C.POINTER(Image), _P, C.POINTER(Lors),
                                         C.POINTER(Corrections), _P, _P]),
    "yrtmi_sensitivity_planned": (C.c_int, [_P, _P, C.POINTER(Image), _P, C.POINTER(Lors),
                                            C.POINTER(Image), _P, _P, _P]),
}


def lib():
    """Load libyrtmi.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"yrtmi native library missing: {LIB_PATH}. Build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` - there "
                "is no CPU fallback.")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    return list(_SIGS)


def check(rc: int):
    if rc != 0:
        msg = lib().yrtmi_last_error().decode(errors="replace")
        raise YrtmiError(f"yrtmi error {rc}: {msg}")
    return rc
