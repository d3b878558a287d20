"""Batches past the 32-bit boundaries: BK5 (FP64 and f32w) and the BK3
adjoint at E = 2^20, BK3 at E = 2^19, where the factor arrays hold more
than 2^31 values (and more than 2^32 bytes), so every kernel's element
index and address arithmetic must hold in 64 bits.  The outputs are
poisoned with NaN and must be finite everywhere; the elements on both
sides of each boundary, the last ones and a random sample are checked
against the FP64 oracle.  The inputs are generated on the device (G
uniform in [0.5, 1.5], u, b uniform in [-1, 1], kappa log-normal); about
50 GB of device memory at the peak.  The adjoint runs at E = 2^20, the
limit of the allocation-free entry point, which must give the same L."""
import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import _lib
from omp_bk_amd import geometry as geo
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1e-13


def _straddling(per_elem, E):
    """Elements whose values cross value index 2^28, 2^29, 2^30 or 2^31
    (bytes 2^31 and 2^32 of 8-byte values, 2^32 and 2^33 of 4-byte ones),
    their neighbours, the last two elements and 24 random ones."""
    out = {E - 1, E - 2}
    for bound in (2 ** 28, 2 ** 29, 2 ** 30, 2 ** 31):
        e = bound // per_elem
        out |= {x for x in (e - 1, e, e + 1) if 0 <= x < E}
    out |= set(np.random.default_rng(E).choice(E, 24, replace=False).tolist())
    return np.array(sorted(out))


def _rows(x, E, idx):
    return x.view(E, -1)[torch.as_tensor(idx, device=DEV)].double().cpu().numpy()


def _free():
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_bk5_past_32_bit_boundaries():
    p, E = 7, 1 << 20
    n = p + 1
    g = torch.Generator(device=DEV).manual_seed(5)
    D = torch.as_tensor(geo.bk5_basis(p), device=DEV)
    G = torch.rand(E * 6 * n ** 3, dtype=torch.float64, device=DEV, generator=g) + 0.5
    u = torch.rand(E * n ** 3, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    assert G.numel() > 2 ** 31
    idx = _straddling(6 * n ** 3, E)
    Dh = D.cpu().numpy()
    v = torch.full_like(u, float("nan"))
    ops.bk5(p, D, G, u, out=v)
    torch.cuda.synchronize()
    assert torch.isfinite(v).all()
    ref = O.bk5(n, Dh, _rows(G, E, idx), _rows(u, E, idx))
    assert O.rel_maxdiff(_rows(v, E, idx).reshape(-1), ref) <= TOL
    # the FP32-storage form on the same batch: the FP64 operator of the
    # rounded inputs
    G32, u32 = G.float(), u.float()
    del G, u
    _free()
    v.fill_(float("nan"))
    ops.bk5(p, D, G32, u32, out=v, form="f32w")
    torch.cuda.synchronize()
    assert torch.isfinite(v).all()
    ref = O.bk5(n, Dh, _rows(G32, E, idx), _rows(u32, E, idx))
    assert O.rel_maxdiff(_rows(v, E, idx).reshape(-1), ref) <= TOL
    del G32, u32, v
    _free()


def test_bk3_past_32_bit_boundaries():
    p, E = 7, 1 << 19
    nm, nq = p + 1, p + 2
    g = torch.Generator(device=DEV).manual_seed(3)
    B, Dm = (torch.as_tensor(a, device=DEV) for a in geo.bk13_bases(p))
    G = torch.rand(E * 6 * nq ** 3, dtype=torch.float64, device=DEV, generator=g) + 0.5
    u = torch.rand(E * nm ** 3, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    assert G.numel() > 2 ** 31
    idx = _straddling(6 * nq ** 3, E)
    v = torch.full_like(u, float("nan"))
    ops.bk3(p, B, Dm, G, u, out=v)
    torch.cuda.synchronize()
    assert torch.isfinite(v).all()
    ref = O.bk3(nq, B.cpu().numpy(), Dm.cpu().numpy(), _rows(G, E, idx), _rows(u, E, idx))
    assert O.rel_maxdiff(_rows(v, E, idx).reshape(-1), ref) <= TOL
    del G, u, v
    _free()


def test_adjoint_past_32_bit_boundaries():
    p, E = 6, 1 << 20
    nm, nq = p + 1, p + 2
    g = torch.Generator(device=DEV).manual_seed(6)
    B, Dm = (torch.as_tensor(a, device=DEV) for a in geo.bk13_bases(p))
    G = torch.rand(E * 6 * nq ** 3, dtype=torch.float64, device=DEV, generator=g) + 0.5
    kappa = torch.exp(0.5 * torch.randn(E * nq ** 3, dtype=torch.float64, device=DEV, generator=g))
    u = torch.rand(E * nm ** 3, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    b = torch.rand(E * nm ** 3, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    assert G.numel() > 2 ** 31
    L, dldu, dldk, loss_e = ops.bk3_adjoint(p, B, Dm, G, kappa, u, b)
    torch.cuda.synchronize()
    assert torch.isfinite(dldu).all() and torch.isfinite(dldk).all() and torch.isfinite(loss_e).all()
    idx = _straddling(6 * nq ** 3, E)
    Lr, _, dldu_r, dldk_r = O.bk3_adjoint(nq, B.cpu().numpy(), Dm.cpu().numpy(), _rows(G, E, idx),
                                          _rows(kappa, E, idx), _rows(u, E, idx), _rows(b, E, idx))
    assert O.rel_maxdiff(_rows(dldu, E, idx).reshape(-1), dldu_r) <= TOL
    assert O.rel_maxdiff(_rows(dldk, E, idx).reshape(-1), dldk_r) <= 1e-12
    assert abs(loss_e[torch.as_tensor(idx, device=DEV)].sum().item() - Lr) <= TOL * Lr
    # the allocation-free entry point at its limit: the same L and gradients
    assert E == (1 << 20)
    dldu2 = torch.full_like(dldu, float("nan"))
    L2 = torch.full((1,), float("nan"), dtype=torch.float64, device=DEV)
    st = torch.cuda.current_stream(DEV).cuda_stream
    _lib.call("bkx_bk3_adjoint_loss_f64", p, E, B.data_ptr(), Dm.data_ptr(), G.data_ptr(), kappa.data_ptr(),
              u.data_ptr(), b.data_ptr(), dldu2.data_ptr(), dldk.data_ptr(), L2.data_ptr(), st)
    torch.cuda.synchronize()
    assert L2.item() == L.item()
    assert torch.equal(dldu2, dldu)
    del G, kappa, u, b, dldu, dldk, loss_e, dldu2
    _free()
