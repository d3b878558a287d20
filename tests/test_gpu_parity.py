"""Parity of the HIP operators with the CPU oracle, through the C ABI.

Bar: max-relative difference max|v - ref| / max|ref| <= 1e-13 in FP64 (the
kernels fuse multiply-adds and reorder the BK5/BK3 divergence sums, so they
match the reference to rounding, not bitwise).  Outputs are poisoned with
NaN before every call so that a missed entry fails.  At the full workload
sizes the oracle checks a sample of elements (the operators are element-
local) and size-independent properties: constants are annihilated, the
operator is symmetric, and the constant-data norm equals the reference
driver's (include/bk_common.h:44-58, openmp/BK3.cpp:50-52, BK5.cpp:143-145).
"""
import glob
import os

import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import geometry as geo
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = 1e-13
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dev(x):
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device=DEV)


def poisoned(n):
    return torch.full((n,), float("nan"), dtype=torch.float64, device=DEV)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().reshape(-1)


def relerr(x, ref):
    assert np.all(np.isfinite(x)), "entries left unwritten (NaN poison) or non-finite"
    return O.rel_maxdiff(x, ref)


def run(kind, p, *arrays):
    u = arrays[-1]
    out = poisoned(u.size)
    fn = {1: ops.bk1, 3: ops.bk3, 5: ops.bk5}[kind]
    fn(p, *[dev(a) for a in arrays], out=out)
    return host(out)


def oracle(kind, p, *arrays):
    if kind == 1:
        return O.bk1(p + 2, *arrays)
    if kind == 3:
        return O.bk3(p + 2, *arrays)
    return O.bk5(p + 1, *arrays)


def lcg_case(kind, p, E, seed):
    if kind == 5:
        n = p + 1
        u, G = O.fill_random(E * n ** 3, E * 6 * n ** 3, seed=seed)
        return (O.test_basis(n, n), G, u)
    nm, nq = p + 1, p + 2
    B = O.test_basis(nm, nq)
    if kind == 1:
        u, J = O.fill_random(E * nm ** 3, E * nq ** 3, seed=seed)
        return (B, J, u)
    u, G = O.fill_random(E * nm ** 3, E * 6 * nq ** 3, seed=seed)
    return (B, O.test_basis(nq, nq), G, u)


def geo_case(kind, p, E, seed):
    rng = np.random.default_rng(seed)
    shape = geo.box_shape(E) if (E & (E - 1)) == 0 else (E, 1, 1)
    if kind == 5:
        return (geo.bk5_basis(p), geo.geometric_factors(5, p, E, shape=shape), rng.uniform(-1, 1, E * (p + 1) ** 3))
    B, D = geo.bk13_bases(p)
    u = rng.uniform(-1, 1, E * (p + 1) ** 3)
    if kind == 1:
        return (B, geo.geometric_factors(1, p, E, shape=shape), u)
    return (B, D, geo.geometric_factors(3, p, E, shape=shape), u)


ORDERS = [(1, p) for p in range(1, 9)] + [(3, p) for p in range(1, 9)] + [(5, p) for p in range(1, 8)]


@pytest.mark.parametrize("kind,p", ORDERS)
def test_golden(kind, p):
    g = np.load(os.path.join(GOLDEN, "bk%d_p%d.npz" % (kind, p)))
    names = {1: ("basis", "JxW", "u"), 3: ("basis", "dbasis", "G", "u"), 5: ("dbasis", "G", "u")}[kind]
    got = run(kind, p, *[g[k] for k in names])
    assert relerr(got, g["sumfact"]) <= TOL
    assert relerr(got, g["harness"]) <= TOL


@pytest.mark.parametrize("E", [1, 2, 7, 61])
@pytest.mark.parametrize("kind,p", ORDERS)
def test_lcg_ragged(kind, p, E):
    """Element counts that leave partial workgroups (and partial wavefronts)."""
    case = lcg_case(kind, p, E, seed=1000 + 17 * E + p)
    assert relerr(run(kind, p, *case), oracle(kind, p, *case)) <= TOL


@pytest.mark.parametrize("kind,p", ORDERS)
def test_real_geometry(kind, p):
    case = geo_case(kind, p, 64, seed=p)
    assert relerr(run(kind, p, *case), oracle(kind, p, *case)) <= TOL


def test_bk1_config_whole_mesh():
    """The BK1 workload itself: p = 3 (n = 4, q = 5), E = 4096 elements of the
    16x16x16 perturbed box mesh, JxW = w detJ in FP64, u uniform[-1, 1] from
    a seed; every element against the oracle (no sampling: 0.26 M DoF)."""
    p, E = 3, 4096
    assert geo.box_shape(E) == (16, 16, 16)
    case = geo_case(1, p, E, seed=4096)
    assert relerr(run(1, p, *case), oracle(1, p, *case)) <= TOL


def test_empty_and_errors():
    z = torch.zeros(0, dtype=torch.float64, device=DEV)
    D = dev(geo.bk5_basis(3))
    assert ops.bk5(3, D, z, z).numel() == 0
    with pytest.raises(ValueError):
        ops.bk5(8, dev(np.zeros(81)), z, z)
    with pytest.raises(TypeError):
        ops.bk5(3, D, z.float(), z)
    # no in-place apply: the kernels read u while other workgroups write out
    E = 3
    u = torch.rand(E * 64, dtype=torch.float64, device=DEV)
    G = torch.rand(E * 6 * 64, dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError):
        ops.bk5(3, D, G, u, out=u)
    B, Dm = (dev(a) for a in geo.bk13_bases(3))
    G3 = torch.rand(E * 6 * 125, dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError):
        ops.bk3(3, B, Dm, G3, u, out=u)
    with pytest.raises(ValueError):
        ops.bk1(3, B, torch.rand(E * 125, dtype=torch.float64, device=DEV), u, out=u)


def _sample_check(kind, p, args_dev, out, nm, nq, sample):
    """Oracle on a sample of elements of a large run."""
    E = out.numel() // nm ** 3
    idx = torch.as_tensor(sample, device=DEV)
    if kind == 5:
        D, G, u = args_dev
        Gs = G.view(E, -1)[idx].cpu().numpy()
        us = u.view(E, -1)[idx].cpu().numpy()
        ref = O.bk5(nq, D.cpu().numpy(), Gs, us)
    else:
        B, D, G, u = args_dev
        Gs = G.view(E, -1)[idx].cpu().numpy()
        us = u.view(E, -1)[idx].cpu().numpy()
        ref = O.bk3(nq, B.cpu().numpy(), D.cpu().numpy(), Gs, us)
    got = out.view(E, -1)[idx].cpu().numpy()
    return relerr(got.reshape(-1), ref)


@pytest.mark.parametrize("kind,p,E", [(5, 7, 65536), (3, 5, 32768), (3, 7, 16384)])
def test_full_size(kind, p, E):
    """Workload sizes: sampled elements against the oracle, plus symmetry and
    the annihilation of constants over the whole mesh."""
    torch.manual_seed(0)
    if kind == 5:
        nm = nq = p + 1
        D = dev(geo.bk5_basis(p))
        G, xyz = geo.geometric_factors(5, p, E, device=DEV, coords=True)
        u = (torch.rand(E * nm ** 3, dtype=torch.float64, device=DEV) * 2 - 1) + geo.smooth_mode(xyz).reshape(-1)
        args = (D, G.reshape(-1), u)
        apply = lambda x: ops.bk5(p, D, G.reshape(-1), x)
    else:
        nm, nq = p + 1, p + 2
        B, Dm = (dev(a) for a in geo.bk13_bases(p))
        G = geo.geometric_factors(3, p, E, device=DEV).reshape(-1)
        u = torch.rand(E * nm ** 3, dtype=torch.float64, device=DEV) * 2 - 1
        args = (B, Dm, G, u)
        apply = lambda x: ops.bk3(p, B, Dm, G, x)
    out = poisoned(u.numel())
    out = ops.bk5(p, *args, out=out) if kind == 5 else ops.bk3(p, *args, out=out)
    torch.cuda.synchronize()
    assert torch.isfinite(out).all()
    sample = np.random.default_rng(1).choice(E, 256, replace=False)
    assert _sample_check(kind, p, args, out, nm, nq, sample) <= TOL
    # symmetry: (w, A u) = (u, A w)
    w = torch.rand_like(u) * 2 - 1
    Aw = apply(w)
    lhs, rhs = torch.dot(w, out).item(), torch.dot(u, Aw).item()
    assert abs(lhs - rhs) <= 1e-11 * abs(torch.dot(u, out).item())
    # constants are annihilated
    A1 = apply(torch.ones_like(u))
    assert A1.abs().max().item() <= 1e-11 * G.abs().max().item()


@pytest.mark.parametrize("kind,p,E,norm", [(3, 5, 32768, 1.11889e08), (5, 7, 65536, 356520.0)])
def test_driver_norm_kat(kind, p, E, norm):
    """The reference drivers' constant data (float cos basis, G = 2, in = 3)
    and their printed norm (float arithmetic, 6 digits)."""
    if kind == 5:
        n = p + 1
        D = dev(O.test_basis(n, n, f32=True))
        G = torch.full((E * 6 * n ** 3,), 2.0, dtype=torch.float64, device=DEV)
        u = torch.full((E * n ** 3,), 3.0, dtype=torch.float64, device=DEV)
        out = ops.bk5(p, D, G, u)
    else:
        nm, nq = p + 1, p + 2
        B = dev(O.test_basis(nm, nq, f32=True))
        D = dev(O.test_basis(nq, nq, f32=True))
        G = torch.full((E * 6 * nq ** 3,), 2.0, dtype=torch.float64, device=DEV)
        u = torch.full((E * nm ** 3,), 3.0, dtype=torch.float64, device=DEV)
        out = ops.bk3(p, B, D, G, u)
    got = torch.linalg.vector_norm(out).item()
    assert abs(got - norm) <= 1e-5 * norm
