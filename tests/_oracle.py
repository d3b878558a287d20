"""ctypes bindings of the CPU checker (oracle/liboracle.so) and of the
reference build (oracle/_ref/libbkref.so).  Test infrastructure only: the
product package never imports this module."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libbkref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_sz = ctypes.c_size_t


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


_oracle = None
_ref = None


def oracle():
    """The plain-C restatement; built on first use if missing."""
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)
        lib = ctypes.CDLL(ORACLE_SO)
        for name in ("bko_rel_maxdiff", "bko_rel_l2", "bko_norm2", "bko_bk3_adjoint", "bko_ff_join"):
            getattr(lib, name).restype = ctypes.c_double
        lib.bko_ff_join.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.bko_ff_split.argtypes = [ctypes.c_double, _fp, _fp]
        for n in ("bko_ff_two_sum", "bko_ff_quick_two_sum", "bko_ff_two_prod"):
            getattr(lib, n).argtypes = [ctypes.c_float, ctypes.c_float, _fp, _fp]
        lib.bko_ff_mul.argtypes = [_fp, _fp, ctypes.c_float, ctypes.c_float]
        lib.bko_ff_mul_add.argtypes = [_fp, _fp] + [ctypes.c_float] * 4
        _oracle = lib
    return _oracle


def ref():
    """The reference's own kernels (None where oracle/_ref was not built)."""
    global _ref
    if _ref is None and os.path.exists(REF_SO):
        lib = ctypes.CDLL(REF_SO)
        lib.bkref_rel_maxdiff.restype = ctypes.c_double
        lib.bkref_ff_join.restype = ctypes.c_double
        lib.bkref_ff_join.argtypes = [ctypes.c_float, ctypes.c_float]
        lib.bkref_ff_split.argtypes = [ctypes.c_double, _fp, _fp]
        for n in ("bkref_ff_two_sum", "bkref_ff_quick_two_sum", "bkref_ff_two_prod"):
            getattr(lib, n).argtypes = [ctypes.c_float, ctypes.c_float, _fp, _fp]
        lib.bkref_ff_mul.argtypes = [_fp, _fp, ctypes.c_float, ctypes.c_float]
        lib.bkref_ff_mul_add.argtypes = [_fp, _fp] + [ctypes.c_float] * 4
        _ref = lib
    return _ref


# ---- oracle wrappers (numpy in, numpy out) ---------------------------------

def test_basis(rows, cols, f32=False):
    m = np.empty(rows * cols)
    fn = oracle().bko_make_test_basis_f32 if f32 else oracle().bko_make_test_basis
    fn(_ptr(m), ctypes.c_int(rows), ctypes.c_int(cols))
    return m.reshape(rows, cols)


def fill_random(n_in, n_jxw, seed=12345):
    a = np.empty(n_in)
    b = np.empty(n_jxw)
    oracle().bko_fill_random(_ptr(a), _sz(n_in), _ptr(b), _sz(n_jxw), ctypes.c_uint32(seed))
    return a, b


def bk1(nq, basis, jxw, u, lib=None, name="bko_bk1_sumfact"):
    nm = nq - 1
    u = _arr(u)
    ne = u.size // nm ** 3
    out = np.empty_like(u)
    getattr(lib or oracle(), name)(ctypes.c_int(nq), _sz(ne), _ptr(_arr(basis)), _ptr(_arr(jxw)), _ptr(u), _ptr(out))
    return out


def bk3(nq, basis, D, G, u, name="bko_bk3_sumfact"):
    nm = nq - 1
    u = _arr(u)
    ne = u.size // nm ** 3
    out = np.empty_like(u)
    getattr(oracle(), name)(ctypes.c_int(nq), _sz(ne), _ptr(_arr(basis)), _ptr(_arr(D)), _ptr(_arr(G)), _ptr(u),
                            _ptr(out))
    return out


def bk3_kappa(nq, basis, D, G, kappa, u):
    nm = nq - 1
    u = _arr(u)
    ne = u.size // nm ** 3
    out = np.empty_like(u)
    oracle().bko_bk3_sumfact_kappa(ctypes.c_int(nq), _sz(ne), _ptr(_arr(basis)), _ptr(_arr(D)), _ptr(_arr(G)),
                                   _ptr(_arr(kappa)), _ptr(u), _ptr(out))
    return out


def bk3_adjoint(nq, basis, D, G, kappa, u, b):
    u = _arr(u)
    dldu = np.empty_like(u)
    r = np.empty_like(u)
    ne = u.size // (nq - 1) ** 3
    dldk = np.empty(ne * nq ** 3)
    loss = oracle().bko_bk3_adjoint(ctypes.c_int(nq), _sz(ne), _ptr(_arr(basis)), _ptr(_arr(D)), _ptr(_arr(G)),
                                    _ptr(_arr(kappa)), _ptr(u), _ptr(_arr(b)), _ptr(r), _ptr(dldu), _ptr(dldk))
    return loss, r, dldu, dldk


def bk5(nq, D, G, u, name="bko_bk5_sumfact"):
    u = _arr(u)
    ne = u.size // nq ** 3
    out = np.empty_like(u)
    getattr(oracle(), name)(ctypes.c_int(nq), _sz(ne), _ptr(_arr(D)), _ptr(_arr(G)), _ptr(u), _ptr(out))
    return out


def bk1_ff(nq, basis, jxw, u):
    return bk1(nq, basis, jxw, u, name="bko_bk1_ff_sumfact")


def rel_maxdiff(x, r):
    x, r = _arr(x), _arr(r)
    return oracle().bko_rel_maxdiff(_ptr(x), _ptr(r), _sz(x.size))


def rel_l2(x, r):
    x, r = _arr(x), _arr(r)
    return oracle().bko_rel_l2(_ptr(x), _ptr(r), _sz(x.size))


# ---- reference wrappers ----------------------------------------------------

def ref_bk1(p, basis, jxw, u):
    nq = p + 2
    u = _arr(u)
    out = np.empty_like(u)
    rc = ref().bkref_bk1_sumfact(ctypes.c_int(p), _sz(u.size // (nq - 1) ** 3), _ptr(_arr(basis)), _ptr(_arr(jxw)),
                                 _ptr(u), _ptr(out))
    assert rc == 0
    return out


def ref_bk3(p, basis, D, G, u):
    nq = p + 2
    u = _arr(u)
    out = np.empty_like(u)
    rc = ref().bkref_bk3_sumfact(ctypes.c_int(p), _sz(u.size // (nq - 1) ** 3), _ptr(_arr(basis)), _ptr(_arr(D)),
                                 _ptr(_arr(G)), _ptr(u), _ptr(out))
    assert rc == 0
    return out


def ref_bk5(p, D, G, u):
    nq = p + 1
    u = _arr(u)
    out = np.empty_like(u)
    rc = ref().bkref_bk5_sumfact(ctypes.c_int(p), _sz(u.size // nq ** 3), _ptr(_arr(D)), _ptr(_arr(G)), _ptr(u),
                                 _ptr(out))
    assert rc == 0
    return out


def ref_harness(kind, nq, *arrays):
    """reference_bk1/3/5 of bk_harness.h (float inputs, double arithmetic)."""
    u = _arr(arrays[-1])
    nm = nq if kind == 5 else nq - 1
    out = np.empty(u.size)
    fn = getattr(ref(), "bkref_reference_bk%d" % kind)
    fn(ctypes.c_int(nq), _sz(u.size // nm ** 3), *[_ptr(_arr(a)) for a in arrays[:-1]], _ptr(u), _ptr(out))
    return out


def ref_bk1_ff(variant, p, basis, jxw, u):
    nq = p + 2
    u = _arr(u)
    out = np.empty_like(u)
    rc = ref().bkref_bk1_ff(ctypes.c_int(variant), ctypes.c_int(p), _sz(u.size // (nq - 1) ** 3), _ptr(_arr(basis)),
                            _ptr(_arr(jxw)), _ptr(u), _ptr(out))
    assert rc == 0
    return out


def make_bk5_G(nq, G6i, F, layout):
    """make_bk5_G of the oracle: G6i(e, i, j, s, k) -> F factors in layout
    0 (factor-major), 1 (interleaved) or 2 (column)."""
    G6i = _arr(G6i)
    ne = G6i.size // (6 * nq ** 3)
    out = np.empty(ne * F * nq ** 3)
    oracle().bko_make_bk5_G(ctypes.c_int(nq), _sz(ne), _ptr(G6i), ctypes.c_int(F), ctypes.c_int(layout), _ptr(out))
    return out


def transpose_jxw(nq, jxw):
    jxw = _arr(jxw)
    out = np.empty_like(jxw)
    oracle().bko_transpose_jxw(ctypes.c_int(nq), _sz(jxw.size // nq ** 3), _ptr(jxw), _ptr(out))
    return out


def ref_make_bk5_G(nq, G6i, F, layout):
    """The reference's make_bk5_G (float inputs: G6i must be float-exact)."""
    G6i = _arr(G6i)
    ne = G6i.size // (6 * nq ** 3)
    out = np.empty(ne * F * nq ** 3)
    ref().bkref_make_bk5_G(_ptr(G6i), _sz(ne), ctypes.c_int(nq), ctypes.c_int(F), ctypes.c_int(layout), _ptr(out))
    return out


def ref_transpose_jxw(nq, jxw):
    jxw = _arr(jxw)
    out = np.empty_like(jxw)
    ref().bkref_transpose_jxw(_ptr(jxw), _sz(jxw.size // nq ** 3), ctypes.c_int(nq), _ptr(out))
    return out
