"""Element restriction on the device (bkx_restrict_*) against the numpy
reference, bit for bit (the scatter sums each node in ascending E-vector
order, as the sequential loop does), on the box mesh and on shuffled
offsets; and the assembled operators R^T A R: against dense assembly on a
small mesh, and at the benchmark size symmetric with the constants in their
null space."""
import numpy as np
import pytest
import torch

import dense
import restrict_ref as Rref
from omp_bk_amd import geometry as geo
from omp_bk_amd import ops
from omp_bk_amd.restriction import Restriction

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _t(x):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV)


@pytest.mark.parametrize("p,shape", [(1, (3, 2, 2)), (3, (4, 3, 2)), (7, (4, 4, 4))])
def test_box_mesh_bitwise(p, shape):
    off, lsize = geo.box_offsets(p, shape)
    R = Restriction(off, lsize)
    rng = np.random.default_rng(p)
    uL = rng.uniform(-1, 1, lsize)
    vE = rng.uniform(-1, 1, off.size)
    assert np.array_equal(R.gather(_t(uL)).cpu().numpy(), Rref.gather(off, uL))
    assert np.array_equal(R.scatter(_t(vE)).cpu().numpy(), Rref.scatter(off, vE, lsize))
    R.close()


def test_shuffled_offsets_bitwise():
    """Unstructured offsets: a random node numbering with unused nodes and
    repeated nodes within an element."""
    rng = np.random.default_rng(11)
    E, es, lsize = 500, 27, 4000
    off = rng.integers(0, lsize, (E, es)).astype(np.int32)
    R = Restriction(off, lsize)
    uL = rng.uniform(-1, 1, lsize)
    vE = rng.uniform(-1, 1, off.size)
    assert np.array_equal(R.gather(_t(uL)).cpu().numpy(), Rref.gather(off, uL))
    assert np.array_equal(R.scatter(_t(vE)).cpu().numpy(), Rref.scatter(off, vE, lsize))
    R.close()


@pytest.mark.parametrize("op", [3, 5])
def test_assembled_operator_against_dense(op):
    p, shape = 2, (2, 2, 1)
    E = int(np.prod(shape))
    off, lsize = geo.box_offsets(p, shape)
    R = Restriction(off, lsize)
    uL = np.random.default_rng(op).uniform(-1, 1, lsize)
    if op == 5:
        G = geo.geometric_factors(5, p, E, shape=shape)
        D = geo.bk5_basis(p)
        mats = [dense.bk5_matrix(D, G[e]) for e in range(E)]
        vL = R.scatter(ops.bk5(p, _t(D), _t(G.reshape(-1)), R.gather(_t(uL))))
    else:
        G = geo.geometric_factors(3, p, E, shape=shape)
        B, D = geo.bk13_bases(p)
        mats = [dense.bk3_matrix(B, D, G[e]) for e in range(E)]
        vL = R.scatter(ops.bk3(p, _t(B), _t(D), _t(G.reshape(-1)), R.gather(_t(uL))))
    ref = Rref.assemble(off, mats, lsize) @ uL
    assert np.abs(vL.cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()
    R.close()


def test_assembled_bk5_benchmark_mesh():
    """BK5 p = 7 on the 64 x 32 x 32 mesh with factors from the geometric
    kernel: R^T A R is symmetric and annihilates the constants."""
    p, E = 7, 65536
    shape = geo.box_shape(E)
    off, lsize = geo.box_offsets(p, shape)
    R = Restriction(off, lsize)
    D = torch.as_tensor(geo.bk5_basis(p), device=DEV)
    G = ops.geo_bk5(p, D, _t(geo.gll(p + 1)[1]), geo.nodal_coords(p, E, device=DEV)).reshape(-1)
    A = lambda x: R.scatter(ops.bk5(p, D, G, R.gather(x)))
    g = torch.Generator(device=DEV).manual_seed(5)
    u = torch.rand(lsize, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    w = torch.rand(lsize, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    Au, Aw = A(u), A(w)
    scale = float(Au.abs().max().item())
    assert abs(float(torch.dot(Au, w).item()) - float(torch.dot(u, Aw).item())) <= 1e-11 * scale * lsize ** 0.5
    assert float(A(torch.ones_like(u)).abs().max().item()) <= 1e-11 * scale
    R.close()


def _fused_case(p, shape, seed):
    E = int(np.prod(shape))
    off, lsize = geo.box_offsets(p, shape)
    rng = np.random.default_rng(seed)
    D = _t(geo.bk5_basis(p))
    G = _t(geo.geometric_factors(5, p, E, shape=shape).reshape(-1))
    uL = _t(rng.uniform(-1, 1, lsize))
    return off, lsize, D, G, uL


@pytest.mark.parametrize("p", range(1, 8))
def test_restricted_bk5_equals_gather_then_bk5(p):
    """The gather fused into BK5's loads gives the bits of the separate
    gather and operator, at every order (its default launch shape), on a
    mesh whose element count no workgroup size divides."""
    off, lsize, D, G, uL = _fused_case(p, (5, 3, 3), p)
    R = Restriction(off, lsize)
    ref = ops.bk5(p, D, G, R.gather(uL))
    got = R.bk5(p, D, G, uL)
    assert torch.equal(got, ref)
    R.close()


@pytest.mark.parametrize("variant", [18, 27, 34, 35, 36, 38, 0])
def test_restricted_bk5_launch_shapes(variant):
    """Each launch shape of the fused path (18, 27, 34-36, 38; others run as
    36), and the 8-byte-aligned G that 34-38 fall back from (to 17's
    shape), against the unfused operator of the same shape."""
    from omp_bk_amd import _lib
    L = _lib.lib()
    p = 5
    off, lsize, D, G, uL = _fused_case(p, (4, 3, 3), 40 + variant)
    R = Restriction(off, lsize)
    prev = L.bkx_tune_get(5, 0, p)
    try:
        shape = variant if variant in (18, 27, 34, 35, 36, 38) else 36
        assert L.bkx_tune_set(5, 0, p, shape) == 0
        ref = ops.bk5(p, D, G, R.gather(uL))
        Gs = torch.empty(G.numel() + 1, dtype=torch.float64, device=DEV)[1:]   # 8-byte aligned, not 16
        Gs.copy_(G)
        ref_s = ops.bk5(p, D, Gs, R.gather(uL))
        assert L.bkx_tune_set(5, 0, p, variant) == 0
        assert torch.equal(R.bk5(p, D, G, uL), ref)
        assert torch.equal(R.bk5(p, D, Gs, uL), ref_s)
    finally:
        L.bkx_tune_set(5, 0, p, prev)
    R.close()


def test_restricted_bk5_shuffled_offsets():
    """Unstructured offsets (random node numbering, repeated nodes)."""
    p = 3
    rng = np.random.default_rng(17)
    E, lsize = 300, 5000
    off = rng.integers(0, lsize, (E, (p + 1) ** 3)).astype(np.int32)
    D = _t(geo.bk5_basis(p))
    G = _t(rng.uniform(0.5, 1.5, E * 6 * (p + 1) ** 3))
    uL = _t(rng.uniform(-1, 1, lsize))
    R = Restriction(off, lsize)
    assert torch.equal(R.bk5(p, D, G, uL), ops.bk5(p, D, G, R.gather(uL)))
    with pytest.raises(ValueError):
        R.bk5(4, _t(geo.bk5_basis(4)), G, uL)
    R.close()
