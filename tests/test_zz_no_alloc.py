"""bkx_bk3_adjoint_loss_f64, the adjoint entry point that allocates nothing
(its per-element losses live in a library-owned device array), called
through the C ABI and compared bit for bit with the pair it replaces,
bkx_bk3_adjoint_f64 followed by bkx_sum_f64 over a caller's per-element
losses.  Two calls in a row, the second with fewer elements and another
order, so that the second reuses what the first left in the array; then
no elements (L = 0) and a count above the limit (rejected on the host).
The loss is also checked against the FP64 oracle at the small size."""
import ctypes

import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import _lib
from omp_bk_amd import geometry as geo

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _case(p, E, seed):
    rng = np.random.default_rng(seed)
    B, D = geo.bk13_bases(p)
    G = geo.geometric_factors(3, p, E, shape=(E, 1, 1)).reshape(-1)
    nm, nq = p + 1, p + 2
    kappa = np.exp(0.5 * rng.standard_normal(E * nq ** 3))
    u = rng.uniform(-1, 1, E * nm ** 3)
    b = rng.uniform(-1, 1, E * nm ** 3)
    return [np.ascontiguousarray(a) for a in (B, D, G, kappa, u, b)]


def _dev(a):
    return torch.as_tensor(a, dtype=torch.float64, device=DEV)


def _poisoned(n):
    return torch.full((n,), float("nan"), dtype=torch.float64, device=DEV)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream(DEV).cuda_stream)


def _both(p, E, host):
    """(L, dldu, dldk) from the allocation-free entry point and from
    bkx_bk3_adjoint_f64 + bkx_sum_f64, all copied to the host."""
    B, D, G, kappa, u, b = (_dev(a) for a in host)
    ins = [x.data_ptr() for x in (B, D, G, kappa, u, b)]
    nm3, nq3 = u.numel() // E, kappa.numel() // E
    outs = {}
    for name in ("noalloc", "pair"):
        dldu, dldk, L = _poisoned(E * nm3), _poisoned(E * nq3), _poisoned(1)
        if name == "noalloc":
            _lib.call("bkx_bk3_adjoint_loss_f64", p, E, *ins, dldu.data_ptr(), dldk.data_ptr(), L.data_ptr(),
                      _stream())
        else:
            loss_e = _poisoned(E)
            _lib.call("bkx_bk3_adjoint_f64", p, E, *ins, dldu.data_ptr(), dldk.data_ptr(), loss_e.data_ptr(),
                      _stream())
            _lib.call("bkx_sum_f64", loss_e.data_ptr(), E, L.data_ptr(), _stream())
        torch.cuda.synchronize()
        outs[name] = (L.cpu().numpy(), dldu.cpu().numpy(), dldk.cpu().numpy())
    return outs["noalloc"], outs["pair"]


def test_no_alloc_adjoint_matches_the_allocating_pair():
    # first call: p = 6 at the adjoint workload's order, 3000 elements
    host = _case(6, 3000, seed=61)
    got, want = _both(6, 3000, host)
    for g, w in zip(got, want):
        assert np.array_equal(g, w), "first call differs from adjoint + sum"
    assert np.isfinite(got[0]).all()
    # second call, smaller and of another order: entries 0..E-1 rewritten,
    # the first call's tail left in the array and not read
    E2 = 37
    host2 = _case(4, E2, seed=43)
    got2, want2 = _both(4, E2, host2)
    for g, w in zip(got2, want2):
        assert np.array_equal(g, w), "second (smaller) call differs from adjoint + sum"
    Lr, _, dldu_r, dldk_r = O.bk3_adjoint(6, *host2)
    assert abs(got2[0][0] - Lr) <= 1e-13 * Lr
    assert O.rel_maxdiff(got2[1], dldu_r) <= 1e-13
    assert O.rel_maxdiff(got2[2], dldk_r) <= 1e-13


def test_no_alloc_adjoint_edges():
    host = _case(3, 2, seed=5)
    B, D, G, kappa, u, b = (_dev(a) for a in host)
    ins = [x.data_ptr() for x in (B, D, G, kappa, u, b)]
    dldu, dldk, L = _poisoned(u.numel()), _poisoned(kappa.numel()), _poisoned(1)
    outs = [dldu.data_ptr(), dldk.data_ptr(), L.data_ptr()]
    # no elements: L = 0, nothing else written
    _lib.call("bkx_bk3_adjoint_loss_f64", 3, 0, *ins, *outs, _stream())
    torch.cuda.synchronize()
    assert L.item() == 0.0
    assert torch.isnan(dldu).all()
    # above the array's limit: rejected before any device work
    lim = 1 << 20
    h = _lib.lib()
    assert h.bkx_bk3_adjoint_loss_f64(3, lim + 1, *ins, *outs, _stream()) == 4
    assert h.bkx_bk3_adjoint_loss_f64(9, 2, *ins, *outs, _stream()) == 1
    assert h.bkx_bk3_adjoint_loss_f64(3, 2, *ins, outs[0], outs[1], None, _stream()) == 2
