"""BK3 with the bases in even-odd form (bkx_bk3_apply_prepared_{f64,f32w}),
on the device, against the CPU oracle on the same inputs (the direct
contraction of the reference's loops) and against the direct kernels: the
GLL-to-Gauss basis and the Gauss collocation derivative of every order, LCG
data, element counts that no workgroup size divides, and each launch shape
of the prepared path.

Bars (max-relative): 1e-13 against the oracle, the FP64 parity bound; the
even-odd sums round differently from the direct ones, so against the direct
kernels the bar is 1e-14, not bit equality.  f32w is compared on the same
FP32-rounded inputs (exact FP64 values)."""
import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import _lib
from omp_bk_amd import geometry as geo
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _t(x, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV).to(dt)


def _case(p, E, seed):
    nm, nq = p + 1, p + 2
    B, D = geo.bk13_bases(p)
    u, G = O.fill_random(E * nm ** 3, E * 6 * nq ** 3, seed=seed)
    return B, D, G, u


def _rel(x, r):
    return np.abs(x - r).max() / np.abs(r).max()


@pytest.mark.parametrize("p", range(1, 9))
def test_prepared_f64_against_oracle_and_direct(p):
    B, D, G, u = _case(p, 777, 300 + p)
    P = ops.bk3_basis(p, _t(B), _t(D))
    print("prepared f64 p=%d E=777" % p, flush=True)
    got = ops.bk3_prepared(p, P, _t(G), _t(u)).cpu().numpy()
    direct = ops.bk3(p, _t(B), _t(D), _t(G), _t(u)).cpu().numpy()
    ref = O.bk3(p + 2, B, D, G, u)
    assert _rel(got, ref) <= 1e-13
    assert _rel(got, direct) <= 1e-14


@pytest.mark.parametrize("p", range(1, 9))
def test_prepared_f32w_against_oracle_and_direct(p):
    B, D, G, u = _case(p, 555, 400 + p)
    G32, u32 = G.astype(np.float32), u.astype(np.float32)
    P = ops.bk3_basis(p, _t(B), _t(D))
    print("prepared f32w p=%d E=555" % p, flush=True)
    got = ops.bk3_prepared(p, P, _t(G32, torch.float32), _t(u32, torch.float32), form="f32w").cpu().numpy()
    direct = ops.bk3(p, _t(B), _t(D), _t(G32, torch.float32), _t(u32, torch.float32), form="f32w").cpu().numpy()
    ref = O.bk3(p + 2, B, D, G32.astype(np.float64), u32.astype(np.float64))
    assert _rel(got, ref) <= 1e-13
    assert _rel(got, direct) <= 1e-14


@pytest.mark.parametrize("form,variants", [("f64", (13, 14, 16, 17, 18, 0)), ("f32w", (4, 6, 8, 9, 10, 11, 0))])
@pytest.mark.parametrize("p", [2, 5, 7])
def test_prepared_launch_shapes(form, variants, p):
    L = _lib.lib()
    fid = 0 if form == "f64" else 2
    B, D, G, u = _case(p, 1001, 500 + p)
    dt = torch.float64 if form == "f64" else torch.float32
    if form == "f32w":
        G, u = G.astype(np.float32).astype(np.float64), u.astype(np.float32).astype(np.float64)
    ref = O.bk3(p + 2, B, D, G, u)
    P = ops.bk3_basis(p, _t(B), _t(D))
    prev = L.bkx_tune_get(3, fid, p)
    try:
        for v in variants:
            assert L.bkx_tune_set(3, fid, p, v) == 0
            print("prepared %s p=%d variant %d" % (form, p, v), flush=True)
            got = ops.bk3_prepared(p, P, _t(G, dt), _t(u, dt), form=form).cpu().numpy()
            assert _rel(got, ref) <= 1e-13, v
    finally:
        L.bkx_tune_set(3, fid, p, prev)


def test_bases_that_are_not_centro_symmetric_are_refused():
    with pytest.raises(ValueError):
        ops.bk3_basis(3, O.test_basis(4, 5), O.test_basis(5, 5), device=DEV)


@pytest.mark.parametrize("form", ["f64", "f32w"])
@pytest.mark.parametrize("p", [1, 3, 6, 8])
def test_prepared_adjoint_against_oracle_and_direct(form, p):
    """The fused adjoint (dL/du, dL/dkappa and L) with the bases in even-odd
    form: against the oracle's adjoint and the direct kernel."""
    nm, nq = p + 1, p + 2
    E = 301
    B, D = geo.bk13_bases(p)
    rng = np.random.default_rng(600 + p)
    u, G = O.fill_random(E * nm ** 3, E * 6 * nq ** 3, seed=700 + p)
    kappa = np.exp(0.5 * rng.standard_normal(E * nq ** 3))
    b = rng.uniform(-1, 1, E * nm ** 3)
    dt = torch.float64 if form == "f64" else torch.float32
    if form == "f32w":
        u, G, kappa, b = (x.astype(np.float32).astype(np.float64) for x in (u, G, kappa, b))
    P = ops.bk3_basis(p, _t(B), _t(D))
    print("prepared adjoint %s p=%d E=%d" % (form, p, E), flush=True)
    L, dldu, dldk, _ = ops.bk3_adjoint_prepared(p, P, _t(G, dt), _t(kappa, dt), _t(u, dt), _t(b, dt), form=form)
    Ld, dldu_d, dldk_d, _ = ops.bk3_adjoint(p, _t(B), _t(D), _t(G, dt), _t(kappa, dt), _t(u, dt), _t(b, dt),
                                            form=form)
    loss, _, ref_u, ref_k = O.bk3_adjoint(nq, B, D, G, kappa, u, b)
    got_u, got_k = dldu.cpu().numpy(), dldk.cpu().numpy()
    assert _rel(got_u, ref_u) <= 1e-13 and _rel(got_k, ref_k) <= 1e-13
    assert abs(float(L.item()) - loss) <= 1e-13 * abs(loss)
    assert _rel(got_u, dldu_d.cpu().numpy()) <= 1e-14 and _rel(got_k, dldk_d.cpu().numpy()) <= 1e-14
