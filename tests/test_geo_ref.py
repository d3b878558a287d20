"""Pins the numpy reference of the geometric factors (tests/geo_ref.py)
without a GPU: closed forms on affine meshes (the exact factors of
omp_bk_amd/geometry.py), exactness on polynomial maps of degree p (the
interpolant reproduces them, so J is their derivative), spectral
convergence to the exact factors of the perturbed mesh, and the mesh volume."""
import numpy as np
import pytest

import geo_ref
from omp_bk_amd import geometry as geo


def _rel(x, r):
    return float(np.abs(x - r).max() / np.abs(r).max())


@pytest.mark.parametrize("kind,p", [(1, 1), (1, 3), (3, 2), (3, 5), (3, 8), (5, 1), (5, 4), (5, 7)])
def test_affine_mesh_closed_form(kind, p):
    E, shape = 8, (2, 2, 2)
    coords = geo.nodal_coords(p, E, amp=0.0, shape=shape)
    exact = geo.geometric_factors(kind, p, E, amp=0.0, shape=shape)
    assert _rel(geo_ref.factors(kind, p, coords), exact) <= 1e-13


def _poly_map(p, pts, seed):
    """x_c(xi) = xi_c + 0.04 * (a random polynomial of degree <= p in each of
    xi_1, xi_2, xi_3), and its exact Jacobian, at the tensor points pts^3."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (3, p + 1, p + 1, p + 1)) * 0.04 / (p + 1) ** 3
    P = np.stack([pts ** k for k in range(p + 1)])                       # (p+1, m)
    dP = np.stack([k * pts ** max(k - 1, 0) for k in range(p + 1)])
    m = pts.size
    xi = np.stack(np.meshgrid(pts, pts, pts, indexing="ij"))           # (3, m, m, m)
    x = xi + np.einsum("cabg,ai,bj,gk->cijk", a, P, P, P)
    J = np.empty((m, m, m, 3, 3))
    J[...] = np.eye(3)
    J[..., :, 0] += np.einsum("cabg,ai,bj,gk->ijkc", a, dP, P, P)
    J[..., :, 1] += np.einsum("cabg,ai,bj,gk->ijkc", a, P, dP, P)
    J[..., :, 2] += np.einsum("cabg,ai,bj,gk->ijkc", a, P, P, dP)
    return x, J


@pytest.mark.parametrize("kind,p", [(3, 1), (3, 4), (3, 8), (5, 2), (5, 7)])
def test_polynomial_map_is_exact(kind, p):
    xl, _ = geo.gll(p + 1)
    coords, _ = _poly_map(p, xl, seed=p)
    pts = xl if kind == 5 else geo.gauss(p + 2)[0]
    _, J = _poly_map(p, pts, seed=p)
    got = geo_ref.jacobian(kind, p, coords[None])[0]
    assert _rel(got, J) <= 1e-13


def test_converges_to_the_exact_perturbed_factors():
    """The interpolant of the smooth perturbation converges spectrally: the
    error against the exact Jacobian's factors falls with p."""
    E, shape = 64, (4, 4, 4)
    errs = []
    for p in (2, 4, 6, 8):
        coords = geo.nodal_coords(p, E, amp=0.1, shape=shape)
        errs.append(_rel(geo_ref.factors(3, p, coords), geo.geometric_factors(3, p, E, amp=0.1, shape=shape)))
    assert errs[-1] < 1e-4 and all(b < a / 4 for a, b in zip(errs, errs[1:])), errs


@pytest.mark.parametrize("p", [2, 5])
def test_volume(p):
    """Sum of JxW = the volume of the mesh: 1 for the unit cube, and the
    perturbation vanishes on the boundary, so the perturbed mesh fills it."""
    E, shape = 27, (3, 3, 3)
    assert abs(geo_ref.factors(1, p, geo.nodal_coords(p, E, amp=0.0, shape=shape)).sum() - 1.0) <= 1e-13
    # the boundary faces are not displaced, and the quadrature integrates
    # the interpolated map's detJ to rounding
    assert abs(geo_ref.factors(1, p, geo.nodal_coords(p, E, amp=0.1, shape=shape)).sum() - 1.0) <= 1e-12
