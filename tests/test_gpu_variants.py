"""Every launch variant of every operator form against the CPU oracle
(the variants are what the autotuner chooses between, so each must be
correct on its own): ragged element counts, LCG data, <= 1e-13."""
import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import _lib
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FORM = {"f64": 0, "f32c": 1, "f32w": 2, "ff": 4}
# max-relative bars: FP64 arithmetic 1e-13; float-float arithmetic 1e-12
BAR = {"f64": 1e-13, "f32w": 1e-13, "f32c": 1e-12, "ff": 1e-12}


def _dev(x, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV).to(dt)


def _cases():
    out = []
    bk13 = ["f64", "f32w", "f32c", "ff"]
    for op, forms, orders in ((1, bk13, range(1, 9)), (3, bk13, range(1, 9)), (5, ["f64", "f32w"], range(1, 8))):
        for form in forms:
            for v in range(_lib.lib().bkx_tune_count(op, FORM[form])):
                for p in orders:
                    out.append((op, form, v, p))
    return out


@pytest.mark.parametrize("op,form,variant,p", _cases())
def test_variant(op, form, variant, p):
    L = _lib.lib()
    f = FORM[form]
    E = 13
    prev = L.bkx_tune_get(op, f, p)
    _lib.call("bkx_tune_set", op, f, p, variant)
    try:
        if op == 5:
            n = p + 1
            u, G = O.fill_random(E * n ** 3, E * 6 * n ** 3, seed=variant * 100 + p)
            D = O.test_basis(n, n)
            if form == "f32w":
                G = G.astype(np.float32).astype(np.float64)
            dt = torch.float64 if form == "f64" else torch.float32
            out = torch.full((u.size,), float("nan"), dtype=torch.float64, device=DEV)
            ops.bk5(p, _dev(D), _dev(G, dt), _dev(u, dt), out=out, form=form)
            ref = O.bk5(n, D, G, u)
        else:
            nm, nq = p + 1, p + 2
            B = O.test_basis(nm, nq)
            Dm = O.test_basis(nq, nq)
            nf = 1 if op == 1 else 6
            u, Q = O.fill_random(E * nm ** 3, E * nf * nq ** 3, seed=variant * 100 + p)
            dt = torch.float32 if form in ("f32c", "f32w") else torch.float64
            if dt == torch.float32:
                # the FP32 forms' data, exactly (LCG factors >= 2 need 25 bits)
                Q = Q.astype(np.float32).astype(np.float64)
                u = u.astype(np.float32).astype(np.float64)
            out = torch.full((u.size,), float("nan"), dtype=torch.float64, device=DEV)
            if op == 1:
                ops.bk1(p, _dev(B), _dev(Q, dt), _dev(u, dt), out=out, form=form)
                ref = O.bk1(nq, B, Q, u)
            else:
                ops.bk3(p, _dev(B), _dev(Dm), _dev(Q, dt), _dev(u, dt), out=out, form=form)
                ref = O.bk3(nq, B, Dm, Q, u)
        got = out.cpu().numpy()
        assert np.all(np.isfinite(got))
        assert O.rel_maxdiff(got, ref) <= BAR[form]
    finally:
        _lib.call("bkx_tune_set", op, f, p, prev)
        assert L.bkx_tune_get(op, f, p) == prev


# The persistent launch variants (each workgroup walks a strided sequence of
# element groups, prefetching the next group's G and u) only loop when there
# are more element groups than resident workgroups, which the ragged cases
# above never reach.  These cases launch 8192 groups (several passes per
# workgroup on any MI355X partition) with device-generated data, compare the
# whole output with variant 0 and sampled elements with the oracle.
PERSISTENT = {(5, "f64"): (30, 31, 32, 33, 39, 40, 41), (5, "f32w"): (30, 31, 32, 33, 39, 40, 41),
              (3, "f64"): (20, 21, 22, 23)}


def _persistent_cases():
    out = []
    for (op, form), variants in PERSISTENT.items():
        orders = range(5, 8) if op == 5 else range(4, 9)
        for v in variants:
            for p in orders:
                out.append((op, form, v, p))
    return out


@pytest.mark.parametrize("op,form,variant,p", _persistent_cases())
def test_persistent_variant_loops(op, form, variant, p):
    L = _lib.lib()
    f = FORM[form]
    dt = torch.float64 if form == "f64" else torch.float32
    gen = torch.Generator(device=DEV).manual_seed(1000 * variant + p)
    if op == 5:
        n = p + 1
        E = 8192 * max(1, 64 // n ** 2)
        D = O.test_basis(n, n)
        G = (torch.rand(E * 6 * n ** 3, generator=gen, device=DEV, dtype=torch.float64) + 1.0).to(dt)
        u = (torch.rand(E * n ** 3, generator=gen, device=DEV, dtype=torch.float64) * 2 - 1).to(dt)
        run = lambda out: ops.bk5(p, _dev(D), G, u, out=out, form=form)
        per_u, per_g = n ** 3, 6 * n ** 3
        ref_of = lambda g, x: O.bk5(n, D, g, x)
    else:
        nm, nq = p + 1, p + 2
        E = 8192 * max(1, 256 // nq ** 2)
        B, Dm = O.test_basis(nm, nq), O.test_basis(nq, nq)
        G = torch.rand(E * 6 * nq ** 3, generator=gen, device=DEV, dtype=torch.float64) + 1.0
        u = torch.rand(E * nm ** 3, generator=gen, device=DEV, dtype=torch.float64) * 2 - 1
        run = lambda out: ops.bk3(p, _dev(B), _dev(Dm), G, u, out=out, form=form)
        per_u, per_g = nm ** 3, 6 * nq ** 3
        ref_of = lambda g, x: O.bk3(nq, B, Dm, g, x)
    prev = L.bkx_tune_get(op, f, p)
    try:
        base = torch.full((u.numel(),), float("nan"), dtype=torch.float64, device=DEV)
        got = torch.full_like(base, float("nan"))
        _lib.call("bkx_tune_set", op, f, p, 0)
        run(base)
        _lib.call("bkx_tune_set", op, f, p, variant)
        run(got)
        torch.cuda.synchronize()
        assert bool(torch.isfinite(got).all())
        err = float(((got - base).abs().max() / base.abs().max()).item())
        assert err <= BAR[form], err
        # oracle on the first, a middle and the last elements
        for e0 in (0, E // 2 + 1, E - 5):
            ne = 5
            g = G[e0 * per_g:(e0 + ne) * per_g].double().cpu().numpy()
            x = u[e0 * per_u:(e0 + ne) * per_u].double().cpu().numpy()
            ref = ref_of(g, x)
            assert O.rel_maxdiff(got[e0 * per_u:(e0 + ne) * per_u].cpu().numpy(), ref) <= BAR[form]
    finally:
        _lib.call("bkx_tune_set", op, f, p, prev)
