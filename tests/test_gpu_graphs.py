"""The operators captured in a hipGraph (torch.cuda.CUDAGraph): the C ABI
launches record no memory-set nodes and no allocations, so every replay of
the captured work gives the eager result bit for bit.  Each graph is
replayed several times, with eager work and a synchronisation between the
replays and its outputs cleared (by a fill kernel) before each, and every
replay is compared with the eager outputs."""
import numpy as np
import pytest
import torch

from omp_bk_amd import geometry as geo
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
REPLAYS = 3


def _t(x):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV)


def _capture_and_replay(fn, outs):
    """fn() writes the tensors outs (allocated before the capture)."""
    fn()
    torch.cuda.synchronize()
    want = [o.clone() for o in outs]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    for r in range(REPLAYS):
        for o in outs:
            o.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for o, w in zip(outs, want):
            assert torch.equal(o, w), "replay %d" % r
        # eager work between the replays
        junk = torch.randn(1 << 16, device=DEV, dtype=torch.float64)
        (junk * 2).sum().item()
    del g


@pytest.mark.parametrize("kind,p,form", [(1, 3, "f64"), (3, 7, "f64"), (3, 5, "f32w"), (5, 7, "f64"),
                                         (5, 7, "f32w"), (5, 7, "f32c"), (5, 7, "oz")])
def test_operator_graph(kind, p, form):
    E = 513
    rng = np.random.default_rng(kind * 10 + p)
    nm = p + 1
    nq = p + 1 if kind == 5 else p + 2
    nf = 1 if kind == 1 else 6
    st = torch.float64 if form == "f64" else torch.float32
    u = _t(rng.uniform(-1, 1, E * nm ** 3)).to(st)
    Q = _t(rng.uniform(0.5, 1.5, E * nf * nq ** 3)).to(st)
    v = torch.empty(E * nm ** 3, dtype=torch.float64, device=DEV)
    if kind == 5:
        D = _t(geo.bk5_basis(p))
        fn = lambda: ops.bk5(p, D, Q, u, out=v, form=form)
    else:
        B, D = (_t(a) for a in geo.bk13_bases(p))
        if kind == 1:
            fn = lambda: ops.bk1(p, B, Q, u, out=v, form=form)
        else:
            fn = lambda: ops.bk3(p, B, D, Q, u, out=v, form=form)
    _capture_and_replay(fn, [v])


@pytest.mark.parametrize("form", ["f64", "f32w", "f32c"])
def test_adjoint_graph(form):
    # the fused adjoint and the fixed-order sum of the loss, as one graph
    p, E = 6, 300
    rng = np.random.default_rng(3)
    nm, nq = p + 1, p + 2
    st = torch.float64 if form == "f64" else torch.float32
    B, D = (_t(a) for a in geo.bk13_bases(p))
    G = _t(geo.geometric_factors(3, p, E, shape=(E, 1, 1)).reshape(-1)).to(st)
    kappa = _t(np.exp(0.5 * rng.standard_normal(E * nq ** 3))).to(st)
    u = _t(rng.uniform(-1, 1, E * nm ** 3)).to(st)
    b = _t(rng.uniform(-1, 1, E * nm ** 3)).to(st)
    res = {}

    def fn():
        res["r"] = ops.bk3_adjoint(p, B, D, G, kappa, u, b, form=form)

    # the graph's outputs are the tensors the captured call allocated: run
    # the capture, then compare its replays with an eager call
    fn()
    torch.cuda.synchronize()
    want = [x.clone() for x in res["r"]]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    outs = list(res["r"])
    for r in range(REPLAYS):
        for o in outs:
            o.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for o, w in zip(outs, want):
            assert torch.equal(o, w), "replay %d" % r
        junk = torch.randn(1 << 16, device=DEV, dtype=torch.float64)
        (junk * 2).sum().item()
    del g
