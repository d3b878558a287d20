"""The geometric-factor kernels (bkx_geo_bk3_f64, bkx_geo_bk5_f64) against
the numpy reference tests/geo_ref.py (pinned in tests/test_geo_ref.py): on
random polynomial-perturbed elements at every order, and at the benchmark
size against sampled elements and the exact factors of the perturbed mesh
(within the interpolation error); and the factors used by the operators
(BK3 and BK5 annihilate constants with them)."""
import numpy as np
import pytest
import torch

import geo_ref
from omp_bk_amd import geometry as geo
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _t(x):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV)


def _rel(x, r):
    return float(np.abs(x - r).max() / np.abs(r).max())


def _coords(p, E, seed):
    """E elements of a box mesh 1/4 wide (absolute coordinates up to 2) under
    an element-wise random smooth map: each coordinate displaced by a random
    polynomial of degree p in the reference coordinates (small enough to
    keep detJ > 0)."""
    rng = np.random.default_rng(seed)
    xl, _ = geo.gll(p + 1)
    xi = np.stack(np.meshgrid(xl, xl, xl, indexing="ij"))               # (3, n, n, n)
    P = np.stack([xl ** k for k in range(p + 1)])                        # (p+1, n)
    out = np.empty((E, 3, p + 1, p + 1, p + 1))
    for e in range(E):
        origin = rng.uniform(0.0, 1.75, 3)
        a = rng.uniform(-1, 1, (3, p + 1, p + 1, p + 1)) * 0.01 / (p + 1) ** 3
        out[e] = origin[:, None, None, None] + 0.125 * (xi + 1.0) + np.einsum("cabg,ai,bj,gk->cijk", a, P, P, P)
    return out


@pytest.mark.parametrize("p", range(1, 9))
def test_gauss_points_against_reference(p):
    E = 7
    c = _coords(p, E, seed=p)
    B, D = geo.bk13_bases(p)
    _, w = geo.gauss(p + 2)
    G = torch.full((E, 6, p + 2, p + 2, p + 2), float("nan"), dtype=torch.float64, device=DEV)
    J = torch.full((E, p + 2, p + 2, p + 2), float("nan"), dtype=torch.float64, device=DEV)
    ops.geo_bk3(p, _t(B), _t(D), _t(w), _t(c), G=G, JxW=J)
    assert _rel(G.cpu().numpy(), geo_ref.factors(3, p, c)) <= 1e-13
    assert _rel(J.cpu().numpy(), geo_ref.factors(1, p, c)) <= 1e-13


@pytest.mark.parametrize("p,E", [(1, 33), (2, 9), (3, 17), (4, 15), (7, 8)])
def test_gauss_points_partial_workgroups(p, E):
    """Several elements per workgroup (16, 4, 8, 7 and 3 at these orders),
    the last workgroup partly filled: every element's factors as computed
    alone (the first element on its own), and the reference."""
    c = _coords(p, E, seed=20 + p)
    B, D = (_t(a) for a in geo.bk13_bases(p))
    w = _t(geo.gauss(p + 2)[1])
    G, J = ops.geo_bk3(p, B, D, w, _t(c))
    assert _rel(G.cpu().numpy(), geo_ref.factors(3, p, c)) <= 1e-13
    G1, J1 = ops.geo_bk3(p, B, D, w, _t(c[-1:]))
    assert torch.equal(G1[0], G[-1]) and torch.equal(J1[0], J[-1])


@pytest.mark.parametrize("p,E", [(1, 33), (2, 9), (3, 17), (4, 5)])
def test_gll_points_partial_workgroups(p, E):
    """Several elements per wavefront at the low orders (16, 7, 4 and 2),
    the last workgroup partly filled: the reference, and the last element
    computed alone."""
    c = _coords(p, E, seed=30 + p)
    D = _t(geo.bk5_basis(p))
    w = _t(geo.gll(p + 1)[1])
    G = ops.geo_bk5(p, D, w, _t(c))
    assert _rel(G.cpu().numpy(), geo_ref.factors(5, p, c)) <= 1e-13
    G1 = ops.geo_bk5(p, D, w, _t(c[-1:]))
    assert torch.equal(G1[0], G[-1])


@pytest.mark.parametrize("p", range(1, 8))
def test_gll_points_against_reference(p):
    E = 7
    c = _coords(p, E, seed=10 + p)
    _, w = geo.gll(p + 1)
    G = ops.geo_bk5(p, _t(geo.bk5_basis(p)), _t(w), _t(c))
    assert _rel(G.cpu().numpy(), geo_ref.factors(5, p, c)) <= 1e-13


def test_benchmark_mesh_bk5():
    """BK5 p = 7 on the 64 x 32 x 32 perturbed mesh: sampled elements
    against the reference, the whole field against the exact factors (the
    interpolation error of the degree-7 map), and BK5 with these factors
    annihilates constants."""
    p, E = 7, 65536
    n = p + 1
    c = geo.nodal_coords(p, E, device=DEV)
    D = torch.as_tensor(geo.bk5_basis(p), device=DEV)
    G = ops.geo_bk5(p, D, _t(geo.gll(n)[1]), c)
    exact = geo.geometric_factors(5, p, E, device=DEV)
    assert float(((G - exact).abs().max() / exact.abs().max()).item()) <= 1e-5
    for e0 in (0, 31000, E - 3):
        ref = geo_ref.factors(5, p, c[e0:e0 + 3].cpu().numpy())
        assert _rel(G[e0:e0 + 3].cpu().numpy(), ref) <= 1e-13
    ones = torch.ones(E * n ** 3, dtype=torch.float64, device=DEV)
    v = ops.bk5(p, D, G.reshape(-1), ones)
    u = torch.rand(E * n ** 3, dtype=torch.float64, device=DEV)
    scale = float(ops.bk5(p, D, G.reshape(-1), u).abs().max().item())
    assert float(v.abs().max().item()) <= 1e-11 * scale


def test_benchmark_mesh_bk3():
    """BK3 p = 5 on the 32^3 perturbed mesh: sampled elements against the
    reference, the field against the exact factors, JxW sums to the volume,
    and BK3 with these factors annihilates constants."""
    p, E = 5, 32768
    nm, nq = p + 1, p + 2
    c = geo.nodal_coords(p, E, device=DEV)
    B, Dm = (torch.as_tensor(a, device=DEV) for a in geo.bk13_bases(p))
    G, J = ops.geo_bk3(p, B, Dm, _t(geo.gauss(nq)[1]), c)
    exact = geo.geometric_factors(3, p, E, device=DEV)
    assert float(((G - exact).abs().max() / exact.abs().max()).item()) <= 1e-4
    assert abs(float(J.sum().item()) - 1.0) <= 1e-11
    for e0 in (0, 17000, E - 3):
        ref = geo_ref.factors(3, p, c[e0:e0 + 3].cpu().numpy())
        assert _rel(G[e0:e0 + 3].cpu().numpy(), ref) <= 1e-13
    ones = torch.ones(E * nm ** 3, dtype=torch.float64, device=DEV)
    v = ops.bk3(p, B, Dm, G.reshape(-1), ones)
    u = torch.rand(E * nm ** 3, dtype=torch.float64, device=DEV)
    scale = float(ops.bk3(p, B, Dm, G.reshape(-1), u).abs().max().item())
    assert float(v.abs().max().item()) <= 1e-11 * scale
