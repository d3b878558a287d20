"""BK3 with a coefficient kappa and its hand-written adjoint, on the device,
against the FP64 oracle (bko_bk3_adjoint: the operator with kappa, the
residual, dL/du = A(kappa) r and dL/dkappa = grad r . G grad u), against
central finite differences of the device loss, and through the
torch.autograd wrapper (gradcheck).  Workload: p = 6, E = 16384, kappa
log-normal (sigma = 0.5), u and b uniform[-1, 1]."""
import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import autograd
from omp_bk_amd import geometry as geo
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _host_case(p, E, seed):
    rng = np.random.default_rng(seed)
    B, D = geo.bk13_bases(p)
    shape = geo.box_shape(E) if (E & (E - 1)) == 0 else (E, 1, 1)
    G = geo.geometric_factors(3, p, E, shape=shape).reshape(-1)
    nm, nq = p + 1, p + 2
    kappa = np.exp(0.5 * rng.standard_normal(E * nq ** 3))
    u = rng.uniform(-1, 1, E * nm ** 3)
    b = rng.uniform(-1, 1, E * nm ** 3)
    return B, D, G, kappa, u, b


def _t(x):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV)


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("E", [1, 7])
def test_against_oracle(p, E):
    B, D, G, kappa, u, b = _host_case(p, E, seed=10 * p + E)
    nq = p + 2
    Lr, r, dldu_r, dldk_r = O.bk3_adjoint(nq, B, D, G, kappa, u, b)
    v = ops.bk3_kappa(p, _t(B), _t(D), _t(G), _t(kappa), _t(u)).cpu().numpy()
    assert O.rel_maxdiff(v, O.bk3_kappa(nq, B, D, G, kappa, u)) <= 1e-13
    L, dldu, dldk, loss_e = ops.bk3_adjoint(p, _t(B), _t(D), _t(G), _t(kappa), _t(u), _t(b))
    assert abs(L.item() - Lr) <= 1e-13 * Lr
    assert O.rel_maxdiff(dldu.cpu().numpy(), dldu_r) <= 1e-13
    assert O.rel_maxdiff(dldk.cpu().numpy(), dldk_r) <= 1e-13
    # VJP with cotangent w: b' = A u - w makes the residual w
    w = np.random.default_rng(p).uniform(-1, 1, u.size)
    _, _, gu_r, gk_r = O.bk3_adjoint(nq, B, D, G, kappa, u, O.bk3_kappa(nq, B, D, G, kappa, u) - w)
    gu, gk = ops.bk3_vjp(p, _t(B), _t(D), _t(G), _t(kappa), _t(u), _t(w))
    assert O.rel_maxdiff(gu.cpu().numpy(), gu_r) <= 1e-13
    assert O.rel_maxdiff(gk.cpu().numpy(), gk_r) <= 1e-12


def _device_case(p, E, seed):
    g = torch.Generator(device=DEV).manual_seed(seed)
    B, D = (torch.as_tensor(a, device=DEV) for a in geo.bk13_bases(p))
    G = geo.geometric_factors(3, p, E, device=DEV).reshape(-1)
    nm, nq = p + 1, p + 2
    kappa = torch.exp(0.5 * torch.randn(E * nq ** 3, dtype=torch.float64, device=DEV, generator=g))
    u = torch.rand(E * nm ** 3, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    b = torch.rand(E * nm ** 3, dtype=torch.float64, device=DEV, generator=g) * 2 - 1
    return B, D, G, kappa, u, b, g


def test_full_size_finite_differences():
    """p = 6, E = 16384: central differences (h = 1e-6) of the device loss
    in 64 random directions each for u and kappa.  L is quadratic in u and
    in kappa separately, so the central difference is exact up to rounding;
    the per-element losses of the two evaluations are differenced before
    they are summed, which keeps that rounding near eps |L_e| sqrt(E) / h."""
    p, E = 6, 16384
    B, D, G, kappa, u, b, g = _device_case(p, E, 7)
    L, dldu, dldk, loss_e = ops.bk3_adjoint(p, B, D, G, kappa, u, b)
    full = 0.5 * torch.sum((ops.bk3_kappa(p, B, D, G, kappa, u) - b) ** 2).item()
    assert abs(full - L.item()) <= 1e-12 * L.item()
    per_elem = lambda uu, kk: ops.bk3_adjoint(p, B, D, G, kk, uu, b)[3]
    h = 1e-6
    # rounding of the differenced per-element losses: a few eps |L_e| per
    # element, summed over E elements as a random walk, divided by 2h
    tol = 8 * 2.2e-16 * loss_e.abs().max().item() * E ** 0.5 / (2 * h)
    for t in range(64):
        du = torch.randn(u.shape, dtype=torch.float64, device=DEV, generator=g)
        dk = torch.randn(kappa.shape, dtype=torch.float64, device=DEV, generator=g)
        fd_u = torch.sum(per_elem(u + h * du, kappa) - per_elem(u - h * du, kappa)).item() / (2 * h)
        fd_k = torch.sum(per_elem(u, kappa + h * dk) - per_elem(u, kappa - h * dk)).item() / (2 * h)
        an_u = torch.dot(dldu, du).item()
        an_k = torch.dot(dldk, dk).item()
        assert abs(fd_u - an_u) <= tol + 1e-9 * abs(an_u), (t, fd_u, an_u, tol)
        assert abs(fd_k - an_k) <= tol + 1e-9 * abs(an_k), (t, fd_k, an_k, tol)


def test_full_size_sampled_oracle():
    """p = 6, E = 16384: a sample of elements against the FP64 oracle."""
    p, E = 6, 16384
    B, D, G, kappa, u, b, _ = _device_case(p, E, 11)
    L, dldu, dldk, loss_e = ops.bk3_adjoint(p, B, D, G, kappa, u, b)
    nm, nq = p + 1, p + 2
    idx = torch.as_tensor(np.random.default_rng(0).choice(E, 128, replace=False), device=DEV)
    h = lambda t, n: t.view(E, n)[idx].cpu().numpy()
    Lr, _, dldu_r, dldk_r = O.bk3_adjoint(nq, B.cpu().numpy(), D.cpu().numpy(), h(G, 6 * nq ** 3),
                                          h(kappa, nq ** 3), h(u, nm ** 3), h(b, nm ** 3))
    assert O.rel_maxdiff(h(dldu, nm ** 3).reshape(-1), dldu_r) <= 1e-13
    assert O.rel_maxdiff(h(dldk, nq ** 3).reshape(-1), dldk_r) <= 1e-13
    assert abs(loss_e[idx].sum().item() - Lr) <= 1e-13 * Lr


def test_autograd_gradcheck():
    p, E = 3, 3
    B, D, G, kappa, u, b, _ = _device_case(p, 8, 5)
    nm, nq = p + 1, p + 2
    u = u[: E * nm ** 3].clone().requires_grad_(True)
    kappa = kappa[: E * nq ** 3].clone().requires_grad_(True)
    G = G[: E * 6 * nq ** 3].contiguous()
    f = lambda uu, kk: autograd.bk3_kappa(uu, kk, p, B, D, G)
    assert torch.autograd.gradcheck(f, (u, kappa), eps=1e-6, atol=1e-7, rtol=1e-6)


def test_autograd_loss_matches_adjoint():
    p, E = 6, 64
    B, D, G, kappa, u, b, _ = _device_case(p, E, 9)
    uu = u.clone().requires_grad_(True)
    kk = kappa.clone().requires_grad_(True)
    v = autograd.bk3_kappa(uu, kk, p, B, D, G)
    L = 0.5 * torch.sum((v - b) ** 2)
    L.backward()
    La, dldu, dldk, _ = ops.bk3_adjoint(p, B, D, G, kappa, u, b)
    assert abs(L.item() - La.item()) <= 1e-13 * La.item()
    assert ((uu.grad - dldu).abs().max() / dldu.abs().max()).item() <= 1e-13
    assert ((kk.grad - dldk).abs().max() / dldk.abs().max()).item() <= 1e-13


@pytest.mark.parametrize("p", range(1, 9))
def test_f32w_against_oracle(p):
    """FP32-storage forms: u, b (w), kappa and G rounded to FP32 are exact
    FP64 inputs of the oracle, so the FP64 arithmetic of the _f32w entry
    points must match it to the FP64 bar."""
    E = 7
    B, D, G, kappa, u, b = _host_case(p, E, seed=70 + p)
    G, kappa, u, b = (x.astype(np.float32) for x in (G, kappa, u, b))
    d = lambda x: x.astype(np.float64)
    f = lambda x: _t(x).float()
    nq = p + 2
    Lr, r, dldu_r, dldk_r = O.bk3_adjoint(nq, B, D, d(G), d(kappa), d(u), d(b))
    v = ops.bk3_kappa(p, _t(B), _t(D), f(G), f(kappa), f(u), form="f32w").cpu().numpy()
    assert O.rel_maxdiff(v, O.bk3_kappa(nq, B, D, d(G), d(kappa), d(u))) <= 1e-13
    L, dldu, dldk, _ = ops.bk3_adjoint(p, _t(B), _t(D), f(G), f(kappa), f(u), f(b), form="f32w")
    assert dldu.dtype == torch.float64 and dldk.dtype == torch.float64
    assert abs(L.item() - Lr) <= 1e-13 * Lr
    assert O.rel_maxdiff(dldu.cpu().numpy(), dldu_r) <= 1e-13
    assert O.rel_maxdiff(dldk.cpu().numpy(), dldk_r) <= 1e-13
    w = np.random.default_rng(p).uniform(-1, 1, u.size).astype(np.float32)
    _, _, gu_r, gk_r = O.bk3_adjoint(nq, B, D, d(G), d(kappa), d(u), O.bk3_kappa(nq, B, D, d(G), d(kappa), d(u)) - d(w))
    gu, gk = ops.bk3_vjp(p, _t(B), _t(D), f(G), f(kappa), f(u), f(w), form="f32w")
    assert O.rel_maxdiff(gu.cpu().numpy(), gu_r) <= 1e-13
    assert O.rel_maxdiff(gk.cpu().numpy(), gk_r) <= 1e-12


def test_f32w_full_size_matches_f64_on_same_inputs():
    """The adjoint configuration (p = 6, E = 16384) in the FP32-storage form
    against the FP64 entry point on the same FP32-rounded inputs."""
    p, E = 6, 16384
    B, D, G, kappa, u, b, _ = _device_case(p, E, seed=7)
    G, kappa, u, b = (x.float() for x in (G, kappa, u, b))
    L64, dldu64, dldk64, _ = ops.bk3_adjoint(p, B, D, G.double(), kappa.double(), u.double(), b.double())
    L32, dldu32, dldk32, _ = ops.bk3_adjoint(p, B, D, G, kappa, u, b, form="f32w")
    torch.cuda.synchronize()
    rel = lambda x, r: float(((x - r).abs().max() / r.abs().max()).item())
    assert abs(L32.item() - L64.item()) <= 1e-13 * L64.item()
    assert rel(dldu32, dldu64) <= 1e-13
    assert rel(dldk32, dldk64) <= 1e-13


@pytest.mark.parametrize("p", range(1, 9))
def test_f32c_against_oracle(p):
    """FP32-compensated forms (float-float arithmetic on FP32 storage)
    against the FP64 oracle on the same FP32-rounded inputs: bound 1e-12
    relative (the float-float sums carry about 44 significant bits; two
    operator applications and the products of dL/dkappa stay within it)."""
    E = 7
    B, D, G, kappa, u, b = _host_case(p, E, seed=90 + p)
    G, kappa, u, b = (x.astype(np.float32) for x in (G, kappa, u, b))
    d = lambda x: x.astype(np.float64)
    f = lambda x: _t(x).float()
    nq = p + 2
    Lr, r, dldu_r, dldk_r = O.bk3_adjoint(nq, B, D, d(G), d(kappa), d(u), d(b))
    v = ops.bk3_kappa(p, _t(B), _t(D), f(G), f(kappa), f(u), form="f32c").cpu().numpy()
    assert O.rel_maxdiff(v, O.bk3_kappa(nq, B, D, d(G), d(kappa), d(u))) <= 1e-12
    L, dldu, dldk, loss_e = ops.bk3_adjoint(p, _t(B), _t(D), f(G), f(kappa), f(u), f(b), form="f32c")
    assert dldu.dtype == torch.float64 and dldk.dtype == torch.float64
    assert abs(L.item() - Lr) <= 1e-12 * Lr
    assert O.rel_maxdiff(dldu.cpu().numpy(), dldu_r) <= 1e-12
    assert O.rel_maxdiff(dldk.cpu().numpy(), dldk_r) <= 1e-12
    w = np.random.default_rng(p).uniform(-1, 1, u.size).astype(np.float32)
    _, _, gu_r, gk_r = O.bk3_adjoint(nq, B, D, d(G), d(kappa), d(u), O.bk3_kappa(nq, B, D, d(G), d(kappa), d(u)) - d(w))
    gu, gk = ops.bk3_vjp(p, _t(B), _t(D), f(G), f(kappa), f(u), f(w), form="f32c")
    assert O.rel_maxdiff(gu.cpu().numpy(), gu_r) <= 1e-12
    assert O.rel_maxdiff(gk.cpu().numpy(), gk_r) <= 1e-12


@pytest.mark.parametrize("p,E", [(1, 1000), (4, 333), (8, 97)])
def test_f32c_ragged_matches_f64(p, E):
    """Ragged element counts (partial last workgroup) in the f32c form
    against the FP64 entry points on the same FP32-rounded inputs."""
    B, D, G, kappa, u, b = _host_case(p, E, seed=3 * p + E)
    f = lambda x: _t(x.astype(np.float32))
    G32, k32, u32, b32 = f(G), f(kappa), f(u), f(b)
    L64, dldu64, dldk64, le64 = ops.bk3_adjoint(p, _t(B), _t(D), G32.double(), k32.double(), u32.double(), b32.double())
    L32, dldu32, dldk32, le32 = ops.bk3_adjoint(p, _t(B), _t(D), G32, k32, u32, b32, form="f32c")
    rel = lambda x, r: float(((x - r).abs().max() / r.abs().max()).item())
    assert abs(L32.item() - L64.item()) <= 1e-12 * L64.item()
    assert rel(le32, le64) <= 1e-12
    assert rel(dldu32, dldu64) <= 1e-12
    assert rel(dldk32, dldk64) <= 1e-12


def test_f32c_full_size_finite_differences():
    """The adjoint configuration (p = 6, E = 16384) in the FP32-compensated
    form: its gradients, on FP32-rounded inputs, against central
    differences (h = 1e-6, 64 random directions each for u and kappa) of
    the FP64 loss at the same point, and against the FP64 entry point's
    gradients to 1e-12.  The perturbed points are not FP32 values, so the
    differences are taken of the FP64 per-element losses."""
    p, E = 6, 16384
    B, D, G, kappa, u, b, g = _device_case(p, E, 13)
    G, kappa, u, b = (x.float() for x in (G, kappa, u, b))
    L, dldu, dldk, loss_e = ops.bk3_adjoint(p, B, D, G, kappa, u, b, form="f32c")
    Gd, kd, ud, bd = (x.double() for x in (G, kappa, u, b))
    L64, dldu64, dldk64, _ = ops.bk3_adjoint(p, B, D, Gd, kd, ud, bd)
    rel = lambda x, r: float(((x - r).abs().max() / r.abs().max()).item())
    assert abs(L.item() - L64.item()) <= 1e-12 * L64.item()
    assert rel(dldu, dldu64) <= 1e-12
    assert rel(dldk, dldk64) <= 1e-12
    per_elem = lambda uu, kk: ops.bk3_adjoint(p, B, D, Gd, kk, uu, bd)[3]
    h = 1e-6
    tol = 8 * 2.2e-16 * loss_e.abs().max().item() * E ** 0.5 / (2 * h)
    for t in range(64):
        du = torch.randn(u.shape, dtype=torch.float64, device=DEV, generator=g)
        dk = torch.randn(kappa.shape, dtype=torch.float64, device=DEV, generator=g)
        fd_u = torch.sum(per_elem(ud + h * du, kd) - per_elem(ud - h * du, kd)).item() / (2 * h)
        fd_k = torch.sum(per_elem(ud, kd + h * dk) - per_elem(ud, kd - h * dk)).item() / (2 * h)
        an_u = torch.dot(dldu, du).item()
        an_k = torch.dot(dldk, dk).item()
        assert abs(fd_u - an_u) <= tol + 1e-9 * abs(an_u), (t, fd_u, an_u, tol)
        assert abs(fd_k - an_k) <= tol + 1e-9 * abs(an_k), (t, fd_k, an_k, tol)
