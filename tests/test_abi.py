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
    # the allocation-free adjoint: order, NULL and its element limit are
    # checked on the host (the fake pointers are never dereferenced)
    fake = [ctypes.c_void_p(256)] * 8
    assert L.bkx_bk3_adjoint_loss_f64(9, 2, *fake, fake[0], None) == 1
    assert L.bkx_bk3_adjoint_loss_f64(3, 2, *fake, None, None) == 2
    assert L.bkx_bk3_adjoint_loss_f64(3, (1 << 20) + 1, *fake, fake[0], None) == 4
    # probes: argument checks
    assert L.bkx_probe_fp64_rate(4, 1, fake[0], 256, None) == 5
    assert L.bkx_probe_fp64_rate(0, 1, fake[0], 100, None) == 5
    assert L.bkx_probe_fp64_rate(0, 1, None, 256, None) == 2
    assert L.bkx_probe_fp64_rate(0, 1, None, 0, None) == 0
    assert L.bkx_probe_read(fake[0], 3, fake[0], None) == 5
    assert L.bkx_probe_read(None, 4, fake[0], None) == 2
    assert L.bkx_probe_read(None, 0, None, None) == 0
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
        # elements-per-workgroup divisors other than 1, 2, 4: rejected before any device array
        good = np.array([3, 5], dtype=np.int32)
        pgood = good.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        div = np.array([0, 1, 1, 3, 1, 1, 1, 1, 1], dtype=np.int32)
        assert create(pgood, 2, div.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), None, 0, ctypes.byref(h)) == 5


def test_python_errors():
    with pytest.raises(_lib.BkxError) as ei:
        _lib.call("bkx_bk5_apply_f64", 0, 1, None, None, None, None, None)
    assert ei.value.status == 1


def test_drivers_reject_bad_arguments_without_device():
    """The precision drivers and the C++ shim caller check the order before
    any device work (the reference prints "unsupported polynomial order"
    and exits 1)."""
    from omp_bk_amd import study
    for kind, p in ((1, 9), (3, 0), (5, 8)):
        r = subprocess.run([os.path.join(study.BIN, "BK%d_ff" % kind), str(p), "10", "1"], capture_output=True,
                           text=True, timeout=60)
        assert r.returncode == 1 and "unsupported polynomial order" in r.stderr, r.stderr
    r = subprocess.run([os.path.join(study.BIN, "bk_shim_check"), "5", "8", "1", "/dev/null", "/dev/null"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "unsupported" in r.stderr, r.stderr
    r = subprocess.run([os.path.join(study.BIN, "bk_shim_check")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr


def test_layout_helpers_check_arguments_without_device():
    L = _lib.lib()
    assert L.bkx_bk5_make_G_f64(8, 10, None, 6, 0, None, None) == 1
    assert L.bkx_bk5_make_G_f64(7, 10, None, 7, 0, None, None) == 5
    assert L.bkx_bk5_make_G_f64(7, 10, None, 9, 3, None, None) == 5
    assert L.bkx_bk5_make_G_f64(7, 0, None, 9, 2, None, None) == 0
    assert L.bkx_bk5_interleave_G_f32(7, 10, 9, 2, None, None, None) == 2
    assert L.bkx_bk1_transpose_jxw_f64(9, 10, None, None, None) == 1
    assert L.bkx_bk1_transpose_jxw_f32(8, 10, None, None, None) == 2
    # FP32-compensated adjoint entry points: order and pointers checked first
    assert L.bkx_bk3_adjoint_f32c(9, 10, *([None] * 10)) == 1
    assert L.bkx_bk3_adjoint_f32c(6, 10, *([None] * 10)) == 2
    assert L.bkx_bk3_vjp_f32c(6, 0, *([None] * 9)) == 0
    assert L.bkx_bk3_apply_kappa_f32c(6, 10, *([None] * 7)) == 2


def test_mixed_plan_checks_per_order_arrays():
    """div_by_order and rounds_by_order are indexed by the order (9 entries):
    a shorter array is refused on the host before the C ABI reads it."""
    from omp_bk_amd import mixed
    order = np.array([1, 2, 3], dtype=np.int32)
    with pytest.raises(ValueError):
        mixed.MixedPlan(order, div_by_order=[1] * 8)
    with pytest.raises(ValueError):
        mixed.MixedPlan(order, rounds_by_order=[1] * 10)


def test_oz_checks_arguments_without_device():
    L = _lib.lib()
    assert L.bkx_bk5_apply_oz(4, 10, None, None, None, None, None) == 1
    assert L.bkx_bk5_apply_oz(6, 10, None, None, None, None, None) == 2
    assert L.bkx_bk5_apply_oz(7, 0, None, None, None, None, None) == 0
    assert L.bkx_bk5_apply_oz(7, 10, None, None, None, None, None) == 2


def test_order_shape_checks_without_device():
    from omp_bk_amd import mixed
    M = mixed._bind()
    d, g = ctypes.c_int(), ctypes.c_int()
    assert M.bkx_mixed_plan_set_order_shape(None, 3, 1, 3) == 2
    assert M.bkx_mixed_plan_get_order_shape(None, 3, ctypes.byref(d), ctypes.byref(g)) == 2


def test_bind_checks_arguments_without_device():
    """A bound apply runs the plain wrapper's checks when it is made: host
    tensors and short arrays are refused before any device call."""
    import torch
    from omp_bk_amd import ops
    p = 3
    B = torch.zeros(4 * 5, dtype=torch.float64)
    J = torch.zeros(2 * 125, dtype=torch.float64)
    u = torch.zeros(2 * 64, dtype=torch.float64)
    with pytest.raises(ValueError, match="HIP device"):
        ops.bind_bk1(p, B, J, u, out=torch.zeros_like(u))
    with pytest.raises(ValueError, match="unsupported polynomial order"):
        ops.bind_bk5(8, B, J, u)
