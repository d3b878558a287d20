"""Bound applies (ops.bind_bk1, bind_bk3, bind_bk5): the arguments checked
once, then each call enqueues the same kernel as the plain wrapper, so the
results are the plain wrapper's bit for bit, for every form; a change of an
input's contents between calls is seen by the next call."""
import pytest
import torch

from omp_bk_amd import geometry as geo
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _rand(n, dtype, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return (torch.rand(n, generator=g, dtype=torch.float64) * 2 - 1).to(DEV, dtype)


@pytest.mark.parametrize("form", ["f64", "ff", "f32c", "f32w"])
@pytest.mark.parametrize("op", [1, 3])
def test_bk13_bound_equals_plain(op, form):
    p, E = 4, 37
    n, q = p + 1, p + 2
    sdt = torch.float32 if form in ("f32c", "f32w") else torch.float64
    B, D = (torch.as_tensor(a, device=DEV) for a in geo.bk13_bases(p))
    u = _rand(E * n ** 3, sdt, 1)
    if op == 1:
        J = (_rand(E * q ** 3, torch.float64, 2) + 2.0).to(sdt)
        plain = lambda v: ops.bk1(p, B, J, u, out=v, form=form)
        bound = ops.bind_bk1(p, B, J, u, out=torch.full((u.numel(),), float("nan"), dtype=torch.float64,
                                                        device=DEV), form=form)
    else:
        G = _rand(E * 6 * q ** 3, sdt, 2)
        plain = lambda v: ops.bk3(p, B, D, G, u, out=v, form=form)
        bound = ops.bind_bk3(p, B, D, G, u, out=torch.full((u.numel(),), float("nan"), dtype=torch.float64,
                                                           device=DEV), form=form)
    for rep in range(3):
        ref = torch.full((u.numel(),), float("nan"), dtype=torch.float64, device=DEV)
        plain(ref)
        got = bound()
        torch.cuda.synchronize()
        assert torch.isfinite(ref).all() and torch.equal(got, ref), (op, form, rep)
        u.mul_(0.5).add_(0.25)            # new contents, same storage


@pytest.mark.parametrize("form,p", [("f64", 7), ("f32w", 7), ("f32c", 5), ("oz", 6), ("bf16s", 7)])
def test_bk5_bound_equals_plain(form, p):
    E = 19
    n = p + 1
    sdt = torch.float64 if form == "f64" else torch.float32
    odt = torch.float32 if form == "bf16s" else torch.float64
    D = torch.as_tensor(geo.bk5_basis(p), device=DEV)
    G = _rand(E * 6 * n ** 3, sdt, 3)
    u = _rand(E * n ** 3, sdt, 4)
    bound = ops.bind_bk5(p, D, G, u, out=torch.empty(u.numel(), dtype=odt, device=DEV), form=form)
    for rep in range(3):
        ref = torch.full((u.numel(),), float("nan"), dtype=odt, device=DEV)
        ops.bk5(p, D, G, u, out=ref, form=form)
        got = bound()
        torch.cuda.synchronize()
        assert torch.isfinite(ref).all() and torch.equal(got, ref), (form, rep)
        G.mul_(1.5)


def test_bound_checks_at_bind():
    p = 3
    B, _ = (torch.as_tensor(a, device=DEV) for a in geo.bk13_bases(p))
    u = torch.zeros(10 * 64, dtype=torch.float64, device=DEV)
    J = torch.zeros(10 * 125, dtype=torch.float64, device=DEV)
    with pytest.raises(ValueError):
        ops.bind_bk1(p, B, J[:-1], u)                  # JxW one value short
    with pytest.raises(ValueError):
        ops.bind_bk1(p, B, J, u, out=u)                # in place
    with pytest.raises(ValueError):
        ops.bind_bk1(9, B, J, u)                       # unsupported order
