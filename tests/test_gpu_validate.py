"""The reference's consistency checks, restated for the HIP operators.

* validate.sh (amx/scripts/validate.sh): every order at an element count
  that no workgroup size divides (1234), LCG data against the oracle (the
  check with discriminating power), and the reference drivers' constant
  data (cos basis, in = 3, JxW = 1 or G = 2; openmp/BK*.cpp) through the
  output norm against the oracle's.
* BK1_ff's claim that all float-float kernels sum every output in the same
  index order and give the same result bit for bit (the reference's
  README.md:223-225): the
  launch variants of each float-float form agree bitwise."""
import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import _lib
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
E_VALIDATE = 1234
FORM = {"f64": 0, "f32c": 1, "f32w": 2, "ff": 4}
ORDERS = [(1, p) for p in range(1, 9)] + [(3, p) for p in range(1, 9)] + [(5, p) for p in range(1, 8)]


def _dev(x, dt=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x), device=DEV).to(dt)


def _case(kind, p, E, constant):
    nm = p + 1
    nq = p + 1 if kind == 5 else p + 2
    nf = {1: 1, 3: 6, 5: 6}[kind]
    if constant:
        u = np.full(E * nm ** 3, 3.0)
        Q = np.full(E * nf * nq ** 3, 1.0 if kind == 1 else 2.0)
    else:
        u, Q = O.fill_random(E * nm ** 3, E * nf * nq ** 3, seed=100 * kind + p)
    if kind == 5:
        return (O.test_basis(nq, nq), Q, u)
    if kind == 1:
        return (O.test_basis(nm, nq), Q, u)
    return (O.test_basis(nm, nq), O.test_basis(nq, nq), Q, u)


def _run(kind, p, arrays, form="f64"):
    dt = torch.float32 if form in ("f32c", "f32w") else torch.float64
    *mats, Q, u = arrays
    out = torch.full((u.size,), float("nan"), dtype=torch.float64, device=DEV)
    fn = {1: ops.bk1, 3: ops.bk3, 5: ops.bk5}[kind]
    fn(p, *[_dev(m) for m in mats], _dev(Q, dt), _dev(u, dt), out=out, form=form)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _oracle(kind, p, arrays):
    nq = p + 1 if kind == 5 else p + 2
    return {1: O.bk1, 3: O.bk3, 5: O.bk5}[kind](nq, *arrays)


@pytest.mark.parametrize("kind,p", ORDERS)
def test_validate_sweep(kind, p):
    rnd = _case(kind, p, E_VALIDATE, constant=False)
    got = _run(kind, p, rnd)
    assert np.all(np.isfinite(got))
    assert O.rel_maxdiff(got, _oracle(kind, p, rnd)) <= 1e-13
    const = _case(kind, p, E_VALIDATE, constant=True)
    n_got = float(np.sqrt(np.sum(_run(kind, p, const) ** 2)))
    n_ref = float(np.sqrt(np.sum(_oracle(kind, p, const) ** 2)))
    assert abs(n_got - n_ref) <= 1e-13 * n_ref


@pytest.mark.parametrize("form", ["ff", "f32c"])
@pytest.mark.parametrize("op", [1, 3])
@pytest.mark.parametrize("p", range(1, 9))
def test_float_float_variants_bitwise(form, op, p):
    L = _lib.lib()
    f = FORM[form]
    count = L.bkx_tune_count(op, f)
    assert count > 1
    arrays = list(_case(op, p, 61, constant=False))
    if form == "f32c":
        arrays[-2:] = [a.astype(np.float32).astype(np.float64) for a in arrays[-2:]]
    prev = L.bkx_tune_get(op, f, p)
    outs = []
    try:
        for v in range(count):
            _lib.call("bkx_tune_set", op, f, p, v)
            outs.append(_run(op, p, arrays, form=form))
    finally:
        _lib.call("bkx_tune_set", op, f, p, prev)
    assert np.all(np.isfinite(outs[0]))
    for v in range(1, count):
        assert np.array_equal(outs[v], outs[0]), "variant %d differs from variant 0" % v
