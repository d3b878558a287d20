"""The layout helpers on the device (make_bk5_G and its inverse,
transpose_jxw; include/bk_harness.h:87-102, 217-244): copies, so the device
results equal the oracle's bit for bit, in FP64 and FP32, at ragged element
counts and at the BK5 workload's size through the round trip."""
import numpy as np
import pytest
import torch

import _oracle as O
from omp_bk_amd import geometry as geo
from omp_bk_amd import ops

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.parametrize("p", range(1, 8))
@pytest.mark.parametrize("E", [1, 13])
def test_make_bk5_G_against_oracle(p, E):
    nq = p + 1
    G6i, _ = O.fill_random(E * 6 * nq ** 3, 1, seed=100 + p + E)
    d = torch.as_tensor(G6i, device=DEV)
    for F in (6, 9):
        for layout in (0, 1, 2):
            ref = O.make_bk5_G(nq, G6i, F, layout)
            got = ops.bk5_make_G(p, d, F, layout)
            assert np.array_equal(got.cpu().numpy(), ref), (F, layout)
            got32 = ops.bk5_make_G(p, d.float(), F, layout)
            assert np.array_equal(got32.cpu().numpy(), ref.astype(np.float32)), (F, layout)
            back = ops.bk5_interleave_G(p, got, F, layout)
            assert torch.equal(back, d), (F, layout)
            assert torch.equal(ops.bk5_interleave_G(p, got32, F, layout), d.float()), (F, layout)


@pytest.mark.parametrize("p", range(1, 9))
@pytest.mark.parametrize("E", [1, 29])
def test_transpose_jxw_against_oracle(p, E):
    nq = p + 2
    jxw = np.random.default_rng(p * E).uniform(0.5, 1.5, E * nq ** 3)
    d = torch.as_tensor(jxw, device=DEV)
    ref = O.transpose_jxw(nq, jxw)
    assert np.array_equal(ops.bk1_transpose_jxw(p, d).cpu().numpy(), ref)
    assert np.array_equal(ops.bk1_transpose_jxw(p, d.float()).cpu().numpy(), ref.astype(np.float32))


def test_factor_major_G_applies_bk5_at_full_size():
    """The BK5 workload's factors (p = 7, E = 65536): nine factors in column
    layout and back gives the same G6i, and BK5 applied to it the same bits."""
    p, E = 7, 65536
    n = p + 1
    G = geo.geometric_factors(5, p, E, device=DEV).reshape(-1)
    D = torch.as_tensor(geo.bk5_basis(p), device=DEV)
    u = torch.rand(E * n ** 3, dtype=torch.float64, device=DEV) * 2 - 1
    G9c = ops.bk5_make_G(p, G, 9, 2)
    back = ops.bk5_interleave_G(p, G9c, 9, 2)
    del G9c
    assert torch.equal(back, G)
    assert torch.equal(ops.bk5(p, D, back, u), ops.bk5(p, D, G, u))


def test_layout_errors():
    z = torch.zeros(6 * 27, dtype=torch.float64, device=DEV)
    with pytest.raises(Exception):
        ops.bk5_make_G(2, z, 7, 0)
    with pytest.raises(Exception):
        ops.bk5_make_G(2, z, 6, 3)
    with pytest.raises(Exception):
        ops.bk5_make_G(8, z, 6, 0)
