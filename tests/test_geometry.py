"""Bases, quadrature and geometric factors pinned by closed forms."""
import numpy as np
import pytest

from omp_bk_amd import geometry as geo


@pytest.mark.parametrize("n", range(2, 11))
def test_gll(n):
    x, w = geo.gll(n)
    assert abs(w.sum() - 2.0) < 1e-14
    assert np.all(np.diff(x) > 0) and x[0] == -1.0 and x[-1] == 1.0
    # exact for polynomials of degree <= 2n - 3
    for d in range(0, 2 * n - 2):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(w @ x ** d - exact) < 1e-13


def test_gll_known_points():
    assert np.allclose(geo.gll(3)[0], [-1, 0, 1], atol=1e-16)
    assert np.allclose(geo.gll(4)[0], [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1], atol=1e-15)
    assert np.allclose(geo.gll(3)[1], [1 / 3, 4 / 3, 1 / 3], atol=1e-15)


@pytest.mark.parametrize("q", range(2, 11))
def test_gauss(q):
    x, w = geo.gauss(q)
    for d in range(0, 2 * q):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(w @ x ** d - exact) < 1e-13


@pytest.mark.parametrize("p", range(1, 9))
def test_lagrange_matrices(p):
    B, D = geo.bk13_bases(p)
    n, q = p + 1, p + 2
    assert B.shape == (n, q) and D.shape == (q, q)
    assert np.allclose(B.sum(axis=0), 1.0, atol=1e-13)          # partition of unity
    assert np.allclose(D.sum(axis=0), 0.0, atol=1e-11)          # derivative of 1
    xg, _ = geo.gauss(q)
    xl, _ = geo.gll(n)
    # interpolation and differentiation are exact for polynomials of their degree
    for d in range(n):
        assert np.allclose(xl ** d @ B, xg ** d, atol=1e-12)
    for d in range(q):
        want = d * xg ** (d - 1) if d else np.zeros(q)
        assert np.allclose(xg ** d @ D, want, atol=1e-10)
    D5 = geo.bk5_basis(p)
    for d in range(n):
        want = d * xl ** (d - 1) if d else np.zeros(n)
        assert np.allclose(D5 @ xl ** d, want, atol=1e-10)


def test_affine_factors():
    """Without displacement an element of an h box has JxW = w h^3 / 8 and
    G = w (h/2) I (rr = ss = tt = w h / 2, cross terms 0)."""
    p, E = 3, 8
    h = 0.5
    J = geo.geometric_factors(1, p, E, amp=0.0)
    G = geo.geometric_factors(3, p, E, amp=0.0)
    _, w = geo.gauss(p + 2)
    w3 = w[:, None, None] * w[None, :, None] * w[None, None, :]
    assert np.allclose(J, w3 * h ** 3 / 8, rtol=1e-15, atol=0)
    for f, diag in enumerate((1, 0, 0, 1, 0, 1)):
        assert np.allclose(G[:, f], diag * w3 * h / 2, rtol=1e-15, atol=1e-18)


def test_perturbed_mesh():
    # volume of the unit cube, G symmetric positive definite, layouts consistent
    J = geo.geometric_factors(1, 4, 512)
    assert abs(J.sum() - 1.0) < 1e-12
    G3 = geo.geometric_factors(3, 4, 64)
    M = np.empty(G3.shape[:1] + G3.shape[2:] + (3, 3))
    for f, (a, b) in enumerate(((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))):
        M[..., a, b] = M[..., b, a] = G3[:, f]
    assert np.linalg.eigvalsh(M).min() > 0
    G5 = geo.geometric_factors(5, 4, 64)
    assert G5.shape == (64, 5, 5, 6, 5)
    x5 = geo.geometric_factors(5, 4, 64, coords=True)[1]
    assert np.all((x5 >= 0) & (x5 <= 1))
    # the displacement vanishes on the cube's faces
    d, _ = geo.displacement(np.array([0.0, 1.0, 0.3]), np.array([0.2, 0.7, 1.0]), np.array([0.5, 0.5, 0.5]),
                            0.1, 1 / 16)
    assert np.allclose(d, 0.0, atol=1e-17)


def test_box_shape():
    assert geo.box_shape(4096) == (16, 16, 16)
    assert geo.box_shape(32768) == (32, 32, 32)
    assert geo.box_shape(65536) == (64, 32, 32)
    with pytest.raises(ValueError):
        geo.box_shape(1000)
