"""Reduced-storage BK5 forms against the FP64 oracle on the same quantised
inputs (u and G rounded to FP32 are exact FP64 values, so the oracle is the
exact operator of the stored data up to FP64 rounding).

Bars: f32c (float-float arithmetic) <= 1e-12 max-relative, the bound the
FP32-compensated form is specified to; f32w (FP64 arithmetic) <= 1e-13.
"""
import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import geometry as geo
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
BARS = {"f32c": 1e-12, "f32w": 1e-13}


def _case(p, E, seed):
    n = p + 1
    rng = np.random.default_rng(seed)
    shape = geo.box_shape(E) if (E & (E - 1)) == 0 else (E, 1, 1)
    G, xyz = geo.geometric_factors(5, p, E, shape=shape, coords=True)
    u = rng.uniform(-1, 1, E * n ** 3) + geo.smooth_mode(xyz).reshape(-1)
    return geo.bk5_basis(p), G.reshape(-1), u


def _run(form, p, D, G, u):
    n = u.size
    out = torch.full((n,), float("nan"), dtype=torch.float64, device=DEV)
    ops.bk5(p, torch.as_tensor(D, device=DEV), torch.as_tensor(G.astype(np.float32), device=DEV),
            torch.as_tensor(u.astype(np.float32), device=DEV), out=out, form=form)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.all(np.isfinite(got))
    return got


@pytest.mark.parametrize("form", ["f32c", "f32w"])
@pytest.mark.parametrize("p", range(1, 8))
@pytest.mark.parametrize("E", [1, 5, 64])
def test_against_quantised_oracle(form, p, E):
    D, G, u = _case(p, E, seed=p * 31 + E)
    Gq = G.astype(np.float32).astype(np.float64)
    uq = u.astype(np.float32).astype(np.float64)
    ref = O.bk5(p + 1, D, Gq, uq)
    got = _run(form, p, D, G, u)
    assert O.rel_maxdiff(got, ref) <= BARS[form]


@pytest.mark.parametrize("form", ["f32c", "f32w"])
def test_lcg_data(form):
    """The reference's LCG data: values that are exact in FP32 already."""
    p, E = 7, 9
    n = p + 1
    u, G = O.fill_random(E * n ** 3, E * 6 * n ** 3, seed=77)
    G = G.astype(np.float32).astype(np.float64)
    D = O.test_basis(n, n)
    got = _run(form, p, D, G, u)
    assert O.rel_maxdiff(got, O.bk5(n, D, G, u)) <= BARS[form]


@pytest.mark.parametrize("form", ["f32c", "f32w"])
def test_full_size_sampled(form):
    """p = 7, E = 65536: sampled elements against the quantised oracle, and
    the error against the unquantised FP64 operator (storage rounding only)."""
    p, E = 7, 65536
    n = p + 1
    D = torch.as_tensor(geo.bk5_basis(p), device=DEV)
    G, xyz = geo.geometric_factors(5, p, E, device=DEV, coords=True)
    G = G.reshape(-1)
    torch.manual_seed(3)
    u = (torch.rand(E * n ** 3, dtype=torch.float64, device=DEV) * 2 - 1) + geo.smooth_mode(xyz).reshape(-1)
    del xyz
    G32, u32 = G.float(), u.float()
    v = ops.bk5(p, D, G32, u32, form=form)
    v64 = ops.bk5(p, D, G, u)
    idx = torch.as_tensor(np.random.default_rng(2).choice(E, 256, replace=False), device=DEV)
    Gs = G32.view(E, -1)[idx].double().cpu().numpy()
    us = u32.view(E, -1)[idx].double().cpu().numpy()
    ref = O.bk5(n, D.cpu().numpy(), Gs, us)
    got = v.view(E, -1)[idx].cpu().numpy().reshape(-1)
    assert O.rel_maxdiff(got, ref) <= BARS[form]
    # against unquantised inputs the error is the FP32 rounding of u and G
    e64 = ((v - v64).abs().max() / v64.abs().max()).item()
    assert 1e-10 < e64 < 1e-5


def test_bf16s_error_and_layout():
    """The split-BF16 MFMA form at p = 7 against the FP64 oracle on the same
    FP32 inputs: a reduced-precision form, bounded at 1e-4 max-relative
    (16-bit operands, FP32 sums); a wrong MFMA operand layout or a missed
    direction gives O(1)."""
    p, E = 7, 37
    D, G, u = _case(p, E, seed=4)
    Gq = G.astype(np.float32).astype(np.float64)
    uq = u.astype(np.float32).astype(np.float64)
    ref = O.bk5(p + 1, D, Gq, uq)
    out = torch.full((u.size,), float("nan"), dtype=torch.float32, device=DEV)
    ops.bk5(p, torch.as_tensor(D, device=DEV), torch.as_tensor(G.astype(np.float32), device=DEV),
            torch.as_tensor(u.astype(np.float32), device=DEV), out=out, form="bf16s")
    got = out.double().cpu().numpy()
    assert np.all(np.isfinite(got))
    err = O.rel_maxdiff(got, ref)
    assert 1e-9 < err < 1e-4, err


def test_bf16s_rejects_other_orders():
    with pytest.raises(Exception):
        ops.bk5(6, torch.zeros(49, dtype=torch.float64, device=DEV), torch.zeros(6 * 343, device=DEV),
                torch.zeros(343, device=DEV), form="bf16s")


def test_bf16s_full_size_sampled():
    """The split-BF16 MFMA form on the full BK5 workload (p = 7, E = 65536,
    64x32x32 perturbed mesh, u = U[-1, 1] + sin(pi x) sin(pi y) sin(pi z)):
    sampled elements against the FP64 oracle on the same FP32 inputs, within
    the form's 1e-4 bound; every output finite."""
    p, E = 7, 65536
    n = p + 1
    D = torch.as_tensor(geo.bk5_basis(p), device=DEV)
    G, xyz = geo.geometric_factors(5, p, E, device=DEV, coords=True)
    G = G.reshape(-1)
    torch.manual_seed(5)
    u = (torch.rand(E * n ** 3, dtype=torch.float64, device=DEV) * 2 - 1) + geo.smooth_mode(xyz).reshape(-1)
    del xyz
    G32, u32 = G.float(), u.float()
    del G, u
    out = torch.full((E * n ** 3,), float("nan"), dtype=torch.float32, device=DEV)
    ops.bk5(p, D, G32, u32, out=out, form="bf16s")
    assert bool(torch.isfinite(out).all().item())
    idx = torch.as_tensor(np.random.default_rng(6).choice(E, 256, replace=False), device=DEV)
    ref = O.bk5(n, D.cpu().numpy(), G32.view(E, -1)[idx].double().cpu().numpy(),
                u32.view(E, -1)[idx].double().cpu().numpy())
    got = out.view(E, -1)[idx].double().cpu().numpy().reshape(-1)
    err = O.rel_maxdiff(got, ref)
    assert 1e-9 < err < 1e-4, err


@pytest.mark.parametrize("p", [5, 6, 7])
@pytest.mark.parametrize("E", [1, 37, 4097])
def test_oz_against_quantised_oracle(p, E):
    """The MFMA-carried compensated form (six exact 8-bit digits per operand,
    BF16 MFMA products summed exactly per digit level, FP64 recombination)
    at p = 5, 6, 7 (n = 6, 7 zero-padded to the 8-deep tiles) against the
    FP64 oracle on the same FP32 inputs, at the FP32-compensated bound of
    1e-12; E = 4097 makes some workgroups loop over two elements."""
    D, G, u = _case(p, E, seed=700 + 10 * p + E)
    ref = O.bk5(p + 1, D, G.astype(np.float32).astype(np.float64), u.astype(np.float32).astype(np.float64))
    got = _run("oz", p, D, G, u)
    err = O.rel_maxdiff(got, ref)
    assert err <= 1e-12, err


@pytest.mark.parametrize("p", [5, 7])
def test_oz_reference_data_and_zero(p):
    """The reference's LCG data on its cos test basis (rows of very different
    size: every row has its own digit exponent), and zero input (exponent 0),
    which must give exactly zero."""
    E = 19
    n = p + 1
    u, G = O.fill_random(E * n ** 3, E * 6 * n ** 3, seed=77)
    G = G.astype(np.float32).astype(np.float64)
    D = O.test_basis(n, n)
    got = _run("oz", p, D, G, u)
    assert O.rel_maxdiff(got, O.bk5(n, D, G, u)) <= 1e-12
    zero = _run("oz", p, D, G, np.zeros_like(u))
    assert not np.any(zero)


def test_oz_full_size_sampled():
    """p = 7, E = 65536 on the 64x32x32 perturbed mesh: sampled elements
    against the quantised oracle at 1e-12, and against f32w (the FP64
    operator on the same FP32 inputs) over the whole output."""
    p, E = 7, 65536
    n = p + 1
    D = torch.as_tensor(geo.bk5_basis(p), device=DEV)
    G, xyz = geo.geometric_factors(5, p, E, device=DEV, coords=True)
    G = G.reshape(-1)
    torch.manual_seed(8)
    u = (torch.rand(E * n ** 3, dtype=torch.float64, device=DEV) * 2 - 1) + geo.smooth_mode(xyz).reshape(-1)
    del xyz
    G32, u32 = G.float(), u.float()
    del G, u
    v = ops.bk5(p, D, G32, u32, form="oz")
    w = ops.bk5(p, D, G32, u32, form="f32w")
    assert ((v - w).abs().max() / w.abs().max()).item() <= 1e-12
    idx = torch.as_tensor(np.random.default_rng(9).choice(E, 256, replace=False), device=DEV)
    ref = O.bk5(n, D.cpu().numpy(), G32.view(E, -1)[idx].double().cpu().numpy(),
                u32.view(E, -1)[idx].double().cpu().numpy())
    assert O.rel_maxdiff(v.view(E, -1)[idx].cpu().numpy().reshape(-1), ref) <= 1e-12


def test_oz_rejects_other_orders():
    with pytest.raises(Exception):
        ops.bk5(4, torch.zeros(25, dtype=torch.float64, device=DEV), torch.zeros(6 * 125, device=DEV),
                torch.zeros(125, device=DEV), form="oz")


@pytest.mark.parametrize("form", ["oz", "bf16s"])
@pytest.mark.parametrize("E", [37, 4097])
def test_unaligned_G_same_bits(form, E):
    """At p = 7 the oz and bf16s kernels read G as 16-byte vectors when G is
    16-byte aligned, and point by point otherwise: G at an address that is
    4-byte but not 16-byte aligned gives the same bits as the aligned copy.
    E = 4097 makes workgroups loop over elements."""
    p = 7
    D, G, u = _case(p, E, seed=900 + E)
    Dt = torch.as_tensor(D, device=DEV)
    Ga = torch.as_tensor(G.astype(np.float32), device=DEV)
    Gs = torch.empty(Ga.numel() + 1, dtype=torch.float32, device=DEV)[1:]
    Gs.copy_(Ga)
    assert Ga.data_ptr() % 16 == 0 and Gs.data_ptr() % 16 != 0
    u32 = torch.as_tensor(u.astype(np.float32), device=DEV)
    odt = torch.float32 if form == "bf16s" else torch.float64
    a = torch.full((u.size,), float("nan"), dtype=odt, device=DEV)
    b = torch.full_like(a, float("nan"))
    ops.bk5(p, Dt, Ga, u32, out=a, form=form)
    ops.bk5(p, Dt, Gs, u32, out=b, form=form)
    assert bool(torch.isfinite(a).all().item())
    assert torch.equal(a, b)
