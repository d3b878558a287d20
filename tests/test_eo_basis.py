"""The even-odd form of centro-symmetric BK3 bases, built on the host by
bkx_bk3_basis_prepare (no device work): its layout reproduces B and D, the
even-odd contraction it encodes equals the direct product, and bases that
are not centro-symmetric (the reference harness's cos test basis) are
refused."""
import ctypes

import numpy as np
import pytest

from omp_bk_amd import _lib
from omp_bk_amd import geometry as geo


def _layout(R, C):
    hr, hc, cr, cc = R // 2, C // 2, (R + 1) // 2, (C + 1) // 2
    cols = 2 * hr * cc + (cc if R % 2 else 0)
    size = cols + 2 * hc * cr + (cr if C % 2 else 0)
    return hr, hc, cr, cc, cols, size


def _prepare(p, B, D):
    L = _lib.lib()
    out = np.full(L.bkx_bk3_basis_size(p), np.nan)
    st = L.bkx_bk3_basis_prepare(p, np.ascontiguousarray(B).ctypes.data, np.ascontiguousarray(D).ctypes.data,
                                 out.ctypes.data)
    return st, out


def _split(p, form):
    nm, nq = p + 1, p + 2
    nb = _layout(nm, nq)[5]
    return form[:nb], form[nb:]


def _rebuild(form, R, C, S):
    """M from the column part of the form (pairs of rows, then the mirror)."""
    hr, hc, cr, cc, cols, size = _layout(R, C)
    P = form[:hr * cc].reshape(hr, cc)
    Q = form[hr * cc:2 * hr * cc].reshape(hr, cc)
    M = np.full((R, C), np.nan)
    M[:hr, :cc] = P + Q
    M[R - 1 - np.arange(hr), :cc] = P - Q
    if R % 2:
        M[R // 2, :cc] = form[2 * hr * cc:cols]
    for c in range(cc, C):                 # M[a][c] = S M[R-1-a][C-1-c]
        M[:, c] = S * M[::-1, C - 1 - c]
    return M


def _eo_cols(form, R, C, S, x):
    """out[c] = sum_m M[m][c] x[m] through the even-odd form."""
    hr, hc, cr, cc, cols, size = _layout(R, C)
    P = form[:hr * cc].reshape(hr, cc)
    Q = form[hr * cc:2 * hr * cc].reshape(hr, cc)
    e = x[:hr] + x[::-1][:hr]
    o = x[:hr] - x[::-1][:hr]
    A = e @ P + (x[R // 2] * form[2 * hr * cc:cols] if R % 2 else 0.0)
    Bq = o @ Q
    out = np.empty(C)
    out[:cc] = A + Bq
    out[C - 1 - np.arange(C // 2)] = S * (A - Bq)[:C // 2]
    return out


def _eo_rows(form, R, C, S, x):
    """out[c] = sum_m M[c][m] x[m] through the even-odd form."""
    hr, hc, cr, cc, cols, size = _layout(R, C)
    P = form[cols:cols + hc * cr].reshape(hc, cr)
    Q = form[cols + hc * cr:cols + 2 * hc * cr].reshape(hc, cr)
    e = x[:hc] + x[::-1][:hc]
    o = x[:hc] - x[::-1][:hc]
    A = e @ P + (x[C // 2] * form[cols + 2 * hc * cr:size] if C % 2 else 0.0)
    Bq = o @ Q
    out = np.empty(R)
    out[:cr] = A + Bq
    out[R - 1 - np.arange(R // 2)] = S * (A - Bq)[:R // 2]
    return out


@pytest.mark.parametrize("p", range(1, 9))
def test_form_of_gll_gauss_bases(p):
    nm, nq = p + 1, p + 2
    B, D = geo.bk13_bases(p)
    st, form = _prepare(p, B, D)
    assert st == 0 and not np.isnan(form).any()
    assert form.size == _layout(nm, nq)[5] + _layout(nq, nq)[5]
    fb, fd = _split(p, form)
    B2, D2 = B.reshape(nm, nq), D.reshape(nq, nq)
    tol = 1e-15 * max(np.abs(B2).max(), np.abs(D2).max())
    assert np.abs(_rebuild(fb, nm, nq, 1) - B2).max() <= 4 * tol
    assert np.abs(_rebuild(fd, nq, nq, -1) - D2).max() <= 64 * tol
    rng = np.random.default_rng(p)
    for M, f, S in ((B2, fb, 1), (D2, fd, -1)):
        R, C = M.shape
        x, y = rng.uniform(-1, 1, R), rng.uniform(-1, 1, C)
        scale = np.abs(M).max() * max(R, C)
        assert np.abs(_eo_cols(f, R, C, S, x) - x @ M).max() <= 1e-14 * scale
        assert np.abs(_eo_rows(f, R, C, S, y) - M @ y).max() <= 1e-14 * scale


def test_sizes_and_errors():
    L = _lib.lib()
    assert L.bkx_bk3_basis_size(7) == 76 + 90   # B 8 x 9: 40 + 36; D 9 x 9: 45 + 45
    assert L.bkx_bk3_basis_size(0) == 0 and L.bkx_bk3_basis_size(9) == 0
    B, D = geo.bk13_bases(3)
    assert _prepare(9, B, D)[0] == 1
    assert L.bkx_bk3_basis_prepare(3, None, None, None) == 2
    # the reference harness's cos test basis is not centro-symmetric
    cosb = np.cos(np.arange(4 * 5, dtype=np.float64))
    cosd = np.cos(np.arange(5 * 5, dtype=np.float64))
    assert _prepare(3, cosb, cosd)[0] == 5
    # a perturbation above the tolerance is refused, one below accepted
    Bp = B.copy().reshape(-1)
    Bp[0] *= 1 + 1e-12
    assert _prepare(3, Bp, D)[0] == 5
    Bp = B.copy().reshape(-1)
    Bp[0] *= 1 + 1e-16
    assert _prepare(3, Bp, D)[0] == 0
    assert L.bkx_bk3_apply_prepared_f64(9, 1, None, None, None, None, None) == 1
    assert L.bkx_bk3_apply_prepared_f64(3, 1, None, None, None, None, None) == 2
    assert L.bkx_bk3_apply_prepared_f32w(3, 0, None, None, None, None, None) == 0


def test_mixed_prepare_matches_per_order_forms():
    from omp_bk_amd import mixed
    L = _lib.lib()
    bases = mixed.packed_bases()
    P = mixed.prepared_bases(bases).numpy()
    assert P.size == sum(L.bkx_bk3_basis_size(p) for p in range(1, 9))
    off, boff = 0, 0
    for p in range(1, 9):
        nm, nq = p + 1, p + 2
        st, form = _prepare(p, bases[boff:boff + nm * nq], bases[boff + nm * nq:boff + nm * nq + nq * nq])
        assert st == 0
        assert np.array_equal(P[off:off + form.size], form)
        off += form.size
        boff += nm * nq + nq * nq
    with pytest.raises(ValueError):
        mixed.prepared_bases(np.cos(np.arange(bases.size, dtype=np.float64)))
