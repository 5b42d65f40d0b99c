"""C ABI: the shared library loads, exports every entry point declared in
include/rectv_amd.h, and fails cleanly (status + message, no crash) where no GPU
or invalid arguments are given.  No compute calls (these run on the CPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rectv_amd.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(rtv_[a-z_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("rtv_create", "rtv_run", "rtv_check_adjoints", "rtv_free", "rtv_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from rectv_gpu_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(declared()) == set(_lib.SIGNATURES), "ctypes signature table out of sync with the header"


def test_failure_paths_report_status_and_message():
    from rectv_gpu_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    # invalid shape is rejected before any device work (nz % nzp != 0)
    st = L.rtv_create(64, 32, 4, 5, 2, 1, ctypes.byref(h))
    assert st == -1 and b"multiple of nzp" in L.rtv_last_error()
    st = L.rtv_create(8, 32, 4, 4, 2, 1, ctypes.byref(h))
    assert st == -1 and b"at least 16" in L.rtv_last_error()
    # a null / freed handle is a state error, not a crash
    assert L.rtv_reset(None) == -4
    assert L.rtv_free(None) == 0
    if _lib.device_count() == 0:
        st = L.rtv_create(64, 32, 4, 4, 2, 1, ctypes.byref(h))
        assert st == -2 and b"device" in L.rtv_last_error().lower()


def test_python_layer_raises_without_device_or_on_bad_shape():
    from rectv_gpu_amd import Solver, _lib
    with pytest.raises(ValueError):
        Solver(64, 32, 4, 5, 2, 1)
    if _lib.device_count() == 0:
        with pytest.raises(_lib.RectvError):
            Solver(64, 32, 4, 4, 2, 1)
