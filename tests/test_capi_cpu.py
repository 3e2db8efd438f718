"""The C-ABI library without a GPU: it loads, exports every entry point
include/yrtmi.h declares, its host-side functions work, and calls that need
a device fail loudly (no silent CPU fallback)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import yrtmi
from helpers import oracle_scanner
from yrtmi import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(REPO, "include", "yrtmi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(yrtmi_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(capi.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/yrtmi.h but not exported"
    # and the ctypes binding covers the whole header
    assert set(names) == set(capi.declared_symbols())


def test_lut_regular_matches_oracle():
    for sc in (yrtmi.Scanner("a", 96.0, 4.0, 4.0, 20.0, 300.0, 256, 24),
               yrtmi.Scanner("b", 200.0, 1.0, 1.0, 10.0, 200.0, 48, 12, 2, 8, 6, 4)):
        pos, ori = sc.lut()
        opos, oori = O.lut(oracle_scanner(sc))
        np.testing.assert_array_equal(pos, opos)
        np.testing.assert_array_equal(ori, oori)


def test_lut_rejects_overlapping_crystals():
    d = capi.ScannerDesc(100.0, 4.0, 40.0, 10.0, 50.0, 64, 4, 1, 1)
    pos = np.zeros((256, 3), np.float32)
    rc = capi.lib().yrtmi_lut_regular(C.byref(d), pos.ctypes.data_as(C.c_void_p),
                                      pos.ctypes.data_as(C.c_void_p))
    assert rc == -1
    assert b"overlap" in capi.lib().yrtmi_last_error()


def test_context_needs_a_device():
    """Without a GPU the context cannot be created and says why."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = capi.lib().yrtmi_ctx_create(0, None, C.byref(h))
    assert rc != 0
    assert capi.lib().yrtmi_last_error() != b""


def test_null_arguments_are_errors():
    L = capi.lib()
    assert L.yrtmi_forward(None, None, None, None, None, None) == -1
    assert L.yrtmi_em_update(None, None, None, None, 1, 1, 1, 0.0) == -1
    assert L.yrtmi_proj_set_launch(None, 4, 256, 0) == -1
    assert L.yrtmi_histo_subset_size(None, 1, 0) == -1
