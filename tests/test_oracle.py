"""The CPU oracle pinned against the reference's golden vectors, the
compiled reference itself (where oracle/_ref was built), dense Kronecker
assembly, invariants of the operators with real bases, the float-float
error-free transformations and finite differences of the adjoint."""
import glob
import os
from fractions import Fraction

import numpy as np
import pytest

import _oracle as O
import dense
from omp_bk_amd import geometry as geo

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(kind):
    files = sorted(glob.glob(os.path.join(GOLDEN, "bk%d_p*.npz" % kind)))
    assert files, "golden vectors missing"
    return [(int(f.rsplit("_p", 1)[1][:-4]), np.load(f)) for f in files]


# ---- golden vectors of the reference (bitwise) -----------------------------

@pytest.mark.parametrize("p,g", _golden(1), ids=lambda x: str(x) if isinstance(x, int) else "")
def test_bk1_golden(p, g):
    nq = p + 2
    assert np.array_equal(O.bk1(nq, g["basis"], g["JxW"], g["u"]), g["sumfact"])
    assert np.array_equal(O.bk1(nq, g["basis"], g["JxW"], g["u"], name="bko_reference_bk1"), g["harness"])
    assert np.array_equal(O.bk1_ff(nq, g["basis"], g["JxW"], g["u"]), g["ff"])


@pytest.mark.parametrize("p,g", _golden(3), ids=lambda x: str(x) if isinstance(x, int) else "")
def test_bk3_golden(p, g):
    nq = p + 2
    assert np.array_equal(O.bk3(nq, g["basis"], g["dbasis"], g["G"], g["u"]), g["sumfact"])
    assert np.array_equal(O.bk3(nq, g["basis"], g["dbasis"], g["G"], g["u"], name="bko_reference_bk3"),
                          g["harness"])
    # the two reference orders agree to rounding only
    assert O.rel_maxdiff(g["sumfact"], g["harness"]) < 1e-13


@pytest.mark.parametrize("p,g", _golden(5), ids=lambda x: str(x) if isinstance(x, int) else "")
def test_bk5_golden(p, g):
    nq = p + 1
    assert np.array_equal(O.bk5(nq, g["dbasis"], g["G"], g["u"]), g["sumfact"])
    assert np.array_equal(O.bk5(nq, g["dbasis"], g["G"], g["u"], name="bko_reference_bk5"), g["harness"])
    assert O.rel_maxdiff(g["sumfact"], g["harness"]) < 1e-13


def test_planted_transpose_is_caught():
    """A transposed D in the oracle (a plausible indexing slip) fails the
    golden comparison by far more than rounding."""
    p, g = _golden(5)[5]
    nq = p + 1
    wrong = O.bk5(nq, g["dbasis"].reshape(nq, nq).T.copy(), g["G"], g["u"])
    assert O.rel_maxdiff(wrong, g["sumfact"]) > 1e-2


def test_test_data_generators():
    # LCG and cos basis against closed forms: first LCG draws by hand
    a, b = O.fill_random(3, 2, seed=12345)
    s = 12345
    want = []
    for _ in range(5):
        s = (1664525 * s + 1013904223) % (1 << 32)
        want.append((s >> 8) / float(1 << 23) - 1.0)
    assert list(a) == want[:3] and list(b) == [1.5 + w for w in want[3:]]
    m = O.test_basis(3, 4)
    assert m[2, 3] == np.cos(11.0)
    ga, gb = geo.lcg_fill(70001, 3)
    oa, ob = O.fill_random(70001, 3)
    assert np.array_equal(ga, oa) and np.array_equal(gb, ob)
    assert np.array_equal(geo.test_basis(7, 9), O.test_basis(7, 9))


@pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (no reference sources)")
@pytest.mark.parametrize("p", range(1, 9))
def test_against_reference_build(p):
    """Fresh random sizes and seeds, bitwise against the compiled reference."""
    nq, nm, E = p + 2, p + 1, 5
    B = O.test_basis(nm, nq)
    D = O.test_basis(nq, nq)
    u, J = O.fill_random(E * nm ** 3, E * nq ** 3, seed=7 + p)
    assert np.array_equal(O.bk1(nq, B, J, u), O.ref_bk1(p, B, J, u))
    u, G = O.fill_random(E * nm ** 3, E * 6 * nq ** 3, seed=99 + p)
    assert np.array_equal(O.bk3(nq, B, D, G, u), O.ref_bk3(p, B, D, G, u))
    for v in range(4):
        assert np.array_equal(O.bk1_ff(nq, B, J, u[: E * nm ** 3]), O.ref_bk1_ff(v, p, B, J, u[: E * nm ** 3]))
    if p <= 7:
        n = p + 1
        Dn = O.test_basis(n, n)
        u, G = O.fill_random(E * n ** 3, E * 6 * n ** 3, seed=3 + p)
        assert np.array_equal(O.bk5(n, Dn, G, u), O.ref_bk5(p, Dn, G, u))


# ---- dense Kronecker assembly ---------------------------------------------

@pytest.mark.parametrize("p", [1, 2, 3, 5])
def test_dense_assembly(p):
    rng = np.random.default_rng(p)
    E = 2
    B, D = geo.bk13_bases(p)
    nm, nq = p + 1, p + 2
    u = rng.uniform(-1, 1, (E, nm ** 3))
    J = geo.geometric_factors(1, p, 8, shape=(2, 2, 2))[:E]
    G = geo.geometric_factors(3, p, 8, shape=(2, 2, 2))[:E]
    v1 = O.bk1(nq, B, J, u).reshape(E, -1)
    v3 = O.bk3(nq, B, D, G, u).reshape(E, -1)
    for e in range(E):
        r1 = dense.bk1_matrix(B, J[e]) @ u[e]
        r3 = dense.bk3_matrix(B, D, G[e]) @ u[e]
        assert np.abs(v1[e] - r1).max() <= 1e-13 * np.abs(r1).max()
        assert np.abs(v3[e] - r3).max() <= 1e-13 * np.abs(r3).max()
    if p <= 7:
        D5 = geo.bk5_basis(p)
        G5 = geo.geometric_factors(5, p, 8, shape=(2, 2, 2))[:E]
        u5 = rng.uniform(-1, 1, (E, (p + 1) ** 3))
        v5 = O.bk5(p + 1, D5, G5, u5).reshape(E, -1)
        for e in range(E):
            r5 = dense.bk5_matrix(D5, G5[e]) @ u5[e]
            assert np.abs(v5[e] - r5).max() <= 1e-13 * np.abs(r5).max()


# ---- invariants with real bases --------------------------------------------

@pytest.mark.parametrize("p", range(1, 8))
def test_invariants(p):
    nm, nq, E = p + 1, p + 2, 8
    B, D = geo.bk13_bases(p)
    J = geo.geometric_factors(1, p, E)
    G = geo.geometric_factors(3, p, E)
    ones = np.ones(E * nm ** 3)
    # partition of unity: 1^T BK1(1) = sum JxW
    assert abs(O.bk1(nq, B, J, ones).sum() - J.sum()) <= 1e-13 * J.sum()
    # derivatives of constants vanish: BK3 and BK5 annihilate constants
    v3 = O.bk3(nq, B, D, G, ones)
    assert np.abs(v3).max() <= 1e-11 * np.abs(G).max()
    D5 = geo.bk5_basis(p)
    G5 = geo.geometric_factors(5, p, E)
    v5 = O.bk5(p + 1, D5, G5, np.ones(E * (p + 1) ** 3))
    assert np.abs(v5).max() <= 1e-11 * np.abs(G5).max()
    # symmetry and semi-definiteness
    rng = np.random.default_rng(10 + p)
    a, b = rng.uniform(-1, 1, (2, E * nm ** 3))
    Aa, Ab = O.bk3(nq, B, D, G, a), O.bk3(nq, B, D, G, b)
    assert abs(a @ Ab - b @ Aa) <= 1e-12 * abs(a @ Aa)
    assert a @ Aa > 0 and b @ Ab > 0
    Ma, Mb = O.bk1(nq, B, J, a), O.bk1(nq, B, J, b)
    assert abs(a @ Mb - b @ Ma) <= 1e-13 * abs(a @ Ma)


@pytest.mark.parametrize("p", [2, 4, 7])
def test_bk5_identity_factors(p):
    """G = identity at every point: BK5 is sum_d D_d^T D_d (einsum)."""
    n, E = p + 1, 3
    D = geo.bk5_basis(p)
    G = np.zeros((E, n, n, 6, n))
    G[:, :, :, 0] = G[:, :, :, 3] = G[:, :, :, 5] = 1.0
    u = np.random.default_rng(p).uniform(-1, 1, (E, n, n, n))
    want = (np.einsum("ni,enjk->eijk", D, np.einsum("nm,emjk->enjk", D, u))
            + np.einsum("nj,eink->eijk", D, np.einsum("nm,eimk->eink", D, u))
            + np.einsum("nk,eijn->eijk", D, np.einsum("nm,eijm->eijn", D, u)))
    got = O.bk5(n, D, G, u).reshape(want.shape)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


# ---- float-float --------------------------------------------------------------

def _f(x):
    return Fraction(float(x))


def test_eft_identities():
    import ctypes
    lib = O.oracle()
    rng = np.random.default_rng(5)
    s, e = ctypes.c_float(), ctypes.c_float()
    for a, b in rng.uniform(-1e3, 1e3, (300, 2)).astype(np.float32):
        lib.bko_ff_two_sum(float(a), float(b), ctypes.byref(s), ctypes.byref(e))
        assert _f(s.value) + _f(e.value) == _f(a) + _f(b)
        lib.bko_ff_two_prod(float(a), float(b), ctypes.byref(s), ctypes.byref(e))
        assert _f(s.value) + _f(e.value) == _f(a) * _f(b)
    hi, lo = ctypes.c_float(), ctypes.c_float()
    for x in rng.uniform(-1, 1, 300):
        lib.bko_ff_split(float(x), ctypes.byref(hi), ctypes.byref(lo))
        assert lib.bko_ff_join(hi.value, lo.value) == np.float64(np.float32(x)) + np.float64(
            np.float32(x - np.float64(np.float32(x))))
        assert abs(lib.bko_ff_join(hi.value, lo.value) - x) <= 2.0 ** -47 * abs(x)


@pytest.mark.skipif(O.ref() is None, reason="oracle/_ref not built (no reference sources)")
def test_eft_bitwise_against_reference():
    import ctypes
    lib, ref = O.oracle(), O.ref()
    rng = np.random.default_rng(6)
    v = rng.uniform(-10, 10, (500, 4)).astype(np.float32)
    for name in ("two_sum", "quick_two_sum", "two_prod"):
        for a, b, _, _ in v:
            x, y, z, w = ctypes.c_float(), ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
            getattr(lib, "bko_ff_" + name)(float(a), float(b), ctypes.byref(x), ctypes.byref(y))
            getattr(ref, "bkref_ff_" + name)(float(a), float(b), ctypes.byref(z), ctypes.byref(w))
            assert (x.value, y.value) == (z.value, w.value)
    for a, b, c, d in v:
        x, y, z, w = (ctypes.c_float(a), ctypes.c_float(b), ctypes.c_float(a), ctypes.c_float(b))
        lib.bko_ff_mul_add(ctypes.byref(x), ctypes.byref(y), float(c), float(d), float(b), float(a))
        ref.bkref_ff_mul_add(ctypes.byref(z), ctypes.byref(w), float(c), float(d), float(b), float(a))
        assert (x.value, y.value) == (z.value, w.value)


@pytest.mark.parametrize("p", range(1, 9))
def test_bk1_ff_accuracy(p):
    """BK1_ff.cpp's check: float-float within 1e-10 (~3e-15) of double."""
    nq, nm, E = p + 2, p + 1, 16
    B = O.test_basis(nm, nq)
    u, J = O.fill_random(E * nm ** 3, E * nq ** 3)
    err = O.rel_l2(O.bk1_ff(nq, B, J, u), O.bk1(nq, B, J, u))
    assert err < 1e-13


def test_kappa_operator_and_metrics():
    """kappa = 1 is BK3 bit for bit; the operator is linear in kappa; the
    norms and error metrics against numpy's closed forms."""
    p, E = 4, 5
    nm, nq = p + 1, p + 2
    B, D = geo.bk13_bases(p)
    G = geo.geometric_factors(3, p, 8)[:E]
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, E * nm ** 3)
    k1, k2 = np.exp(rng.standard_normal((2, E * nq ** 3)))
    assert np.array_equal(O.bk3_kappa(nq, B, D, G, np.ones(E * nq ** 3), u), O.bk3(nq, B, D, G, u))
    lin = O.bk3_kappa(nq, B, D, G, 2.0 * k1 + 3.0 * k2, u)
    ref = 2.0 * O.bk3_kappa(nq, B, D, G, k1, u) + 3.0 * O.bk3_kappa(nq, B, D, G, k2, u)
    assert np.abs(lin - ref).max() <= 1e-13 * np.abs(ref).max()
    x, r = rng.standard_normal((2, 1000))
    import ctypes
    dp = ctypes.POINTER(ctypes.c_double)
    n2 = O.oracle().bko_norm2(x.ctypes.data_as(dp), ctypes.c_size_t(x.size))
    assert abs(n2 - np.linalg.norm(x)) <= 1e-14 * n2
    assert abs(O.rel_l2(x, r) - np.linalg.norm(x - r) / np.linalg.norm(r)) <= 1e-14
    assert O.rel_maxdiff(x, r) == np.abs(x - r).max() / np.abs(r).max()


# ---- adjoint -------------------------------------------------------------------

@pytest.mark.parametrize("p", [2, 5])
def test_adjoint_finite_differences(p):
    nm, nq, E = p + 1, p + 2, 8
    B, D = geo.bk13_bases(p)
    G = geo.geometric_factors(3, p, E)
    rng = np.random.default_rng(p)
    kappa = np.exp(0.5 * rng.standard_normal(E * nq ** 3))
    u, b = rng.uniform(-1, 1, (2, E * nm ** 3))
    L, r, dldu, dldk = O.bk3_adjoint(nq, B, D, G, kappa, u, b)
    assert abs(L - 0.5 * r @ r) <= 1e-14 * L
    h = 1e-6
    for t in range(4):
        du = rng.standard_normal(u.size)
        dk = rng.standard_normal(kappa.size)
        Lp = O.bk3_adjoint(nq, B, D, G, kappa, u + h * du, b)[0]
        Lm = O.bk3_adjoint(nq, B, D, G, kappa, u - h * du, b)[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - dldu @ du) <= 1e-6 * (abs(fd) + 1e-30)
        Lp = O.bk3_adjoint(nq, B, D, G, kappa + h * dk, u, b)[0]
        Lm = O.bk3_adjoint(nq, B, D, G, kappa - h * dk, u, b)[0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - dldk @ dk) <= 1e-6 * (abs(fd) + 1e-30)


@pytest.mark.parametrize("nq", range(2, 9))
def test_layout_helpers(nq):
    """make_bk5_G and transpose_jxw of the oracle: bit for bit against the
    reference build (bk_harness.h:87-102, 217-244), and pinned without it by
    what the layouts fix: layout 1 with six factors is the identity, the
    nine-factor tensor is symmetric with the six factors on and above its
    diagonal, layouts 0 and 2 are the numpy transpositions of layout 1, and
    transposing JxW twice is the identity."""
    E = 3
    G6i, jxw = O.fill_random(E * 6 * nq ** 3, E * nq ** 3, seed=nq)
    jxw = jxw.astype(np.float32).astype(np.float64)      # the reference helpers take floats
    g6 = G6i.reshape(E, nq, nq, 6, nq)                                  # (e, i, j, s, k)
    assert np.array_equal(O.make_bk5_G(nq, G6i, 6, 1), G6i)
    assert np.array_equal(O.make_bk5_G(nq, G6i, 6, 0).reshape(E, 6, nq, nq, nq), g6.transpose(0, 3, 1, 2, 4))
    assert np.array_equal(O.make_bk5_G(nq, G6i, 6, 2).reshape(E, 6, nq, nq, nq), g6.transpose(0, 3, 4, 1, 2))
    g9 = O.make_bk5_G(nq, G6i, 9, 0).reshape(E, 3, 3, nq, nq, nq)       # (e, a, b, i, j, k)
    assert np.array_equal(g9, g9.transpose(0, 2, 1, 3, 4, 5))
    upper = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for s, (a, b) in enumerate(upper):
        assert np.array_equal(g9[:, a, b], g6[:, :, :, s, :])
    t = O.transpose_jxw(nq, jxw)
    assert np.array_equal(t.reshape(E, nq, nq, nq), jxw.reshape(E, nq, nq, nq).transpose(0, 3, 2, 1))
    assert np.array_equal(O.transpose_jxw(nq, t), jxw)
    if O.ref() is None:
        pytest.skip("reference build (oracle/_ref) not present")
    for F in (6, 9):
        for layout in (0, 1, 2):
            assert np.array_equal(O.make_bk5_G(nq, G6i, F, layout), O.ref_make_bk5_G(nq, G6i, F, layout)), (F, layout)
    assert np.array_equal(t, O.ref_transpose_jxw(nq, jxw))
