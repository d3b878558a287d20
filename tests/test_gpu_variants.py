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
