"""The C ABI: libbkx.so loads, exports every function include/bkx.h declares,
and rejects bad arguments before touching the device (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from omp_bk_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(bkx_\w+)\s*\(", txt, re.M))
    return names


def _exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if " T " in line}


def test_every_declared_symbol_is_exported():
    declared = _declared()
    assert "bkx_bk5_apply_f64" in declared
    missing = declared - _exported()
    assert not missing, missing


def test_binding_covers_header():
    assert set(_lib.exported_symbols()) <= _declared()


def test_error_paths_without_device():
    L = _lib.lib()
    assert L.bkx_bk5_apply_f64(0, 10, None, None, None, None, None) == 1
    assert L.bkx_bk5_apply_f64(8, 10, None, None, None, None, None) == 1
    assert L.bkx_bk5_apply_f64(7, 0, None, None, None, None, None) == 0
    assert L.bkx_bk5_apply_f64(7, 10, None, None, None, None, None) == 2
    assert L.bkx_bk3_apply_f64(9, 10, None, None, None, None, None, None) == 1
    assert L.bkx_bk1_apply_f64(3, 10, None, None, None, None, None) == 2
    assert [L.bkx_min_order(k) for k in (1, 3, 5)] == [1, 1, 1]
    assert [L.bkx_max_order(k) for k in (1, 3, 5)] == [8, 8, 7]
    assert L.bkx_max_order(2) == -1
    assert L.bkx_status_string(1) == b"unsupported polynomial order p"
    # geometric factors and element restriction: checked before device work
    assert L.bkx_geo_bk3_f64(9, 10, None, None, None, None, None, None, None) == 1
    assert L.bkx_geo_bk5_f64(8, 10, None, None, None, None, None) == 1
    assert L.bkx_geo_bk5_f64(7, 10, None, None, None, None, None) == 2
    h = ctypes.c_void_p()
    off = np.array([[0, 1, 2, 9]], dtype=np.int32)
    ptr = off.ctypes.data_as(ctypes.c_void_p)
    assert L.bkx_restrict_create(1, 4, None, 10, ctypes.byref(h)) == 2
    assert L.bkx_restrict_create(1, 0, ptr, 10, ctypes.byref(h)) == 5
    assert L.bkx_restrict_create(1, 4, ptr, 9, ctypes.byref(h)) == 5      # offset 9 >= lsize
    assert L.bkx_restrict_gather_f64(None, None, None, None) == 2
    assert L.bkx_restrict_scatter_f64(None, None, None, None) == 2
    assert L.bkx_bk5_apply_restricted_f64(None, 7, None, None, None, None, None) == 2
    # mixed-order plans: checked on the host before any device array is made
    from omp_bk_amd import mixed
    M = mixed._bind()
    bad = np.array([3, 9], dtype=np.int32)
    pbad = bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    for create in (M.bkx_mixed_plan_create, M.bkx_mixed_plan_create_grouped):
        assert create(pbad, 2, None, None, 0, None) == 2
        assert create(None, 2, None, None, 0, ctypes.byref(h)) == 2
        assert create(pbad, 2, None, None, 0, ctypes.byref(h)) == 1


def test_python_errors():
    with pytest.raises(_lib.BkxError) as ei:
        _lib.call("bkx_bk5_apply_f64", 0, 1, None, None, None, None, None)
    assert ei.value.status == 1
