"""Pins the box-mesh restriction offsets (omp_bk_amd.geometry.box_offsets)
and the numpy restriction reference (tests/restrict_ref.py) without a GPU:
node multiplicities, shared nodes with identical coordinates, exact
adjointness, and the assembled operators' null space."""
import numpy as np
import pytest

import dense
import restrict_ref as R
from omp_bk_amd import geometry as geo


@pytest.mark.parametrize("p,shape", [(1, (2, 2, 2)), (2, (3, 2, 1)), (4, (2, 2, 2))])
def test_multiplicities(p, shape):
    off, lsize = geo.box_offsets(p, shape)
    assert lsize == np.prod([s * p + 1 for s in shape])
    m = R.scatter(off, np.ones(off.size), lsize)
    assert m.sum() == off.size and m.min() >= 1
    nx, ny, nz = shape
    grid = m.reshape(nx * p + 1, ny * p + 1, nz * p + 1).transpose(2, 1, 0)   # indexed [gz, gy, gx]

    def share(n, g):   # elements along one axis sharing grid line g
        return 2 if (g % p == 0 and 0 < g < n * p) else 1
    for gz in range(nz * p + 1):
        for gy in range(ny * p + 1):
            for gx in range(nx * p + 1):
                assert grid[gz, gy, gx] == share(nx, gx) * share(ny, gy) * share(nz, gz)


def test_shared_nodes_have_one_position():
    p, shape = 3, (3, 2, 2)
    E = np.prod(shape)
    off, lsize = geo.box_offsets(p, shape)
    xyz = geo.nodal_coords(p, E, amp=0.1, shape=shape).transpose(0, 2, 3, 4, 1).reshape(E * (p + 1) ** 3, 3)
    pos = np.full((lsize, 3), np.nan)
    for x, g in enumerate(off.reshape(-1)):
        if np.isnan(pos[g, 0]):
            pos[g] = xyz[x]
        else:
            assert np.abs(pos[g] - xyz[x]).max() <= 1e-15
    assert not np.isnan(pos).any()


def test_adjoint_exact():
    rng = np.random.default_rng(3)
    off, lsize = geo.box_offsets(2, (2, 3, 2))
    u = rng.integers(-50, 50, lsize).astype(np.float64)
    w = rng.integers(-50, 50, off.size).astype(np.float64)
    assert R.gather(off, u) @ w == u @ R.scatter(off, w, lsize)


def test_assembled_bk5_null_space():
    """The assembled BK5 matrix of a perturbed mesh annihilates the
    constants of the continuous space and is symmetric."""
    p, shape = 2, (2, 2, 1)
    E = int(np.prod(shape))
    off, lsize = geo.box_offsets(p, shape)
    G = geo.geometric_factors(5, p, E, shape=shape)
    D = geo.bk5_basis(p)
    A = R.assemble(off, [dense.bk5_matrix(D, G[e]) for e in range(E)], lsize)
    assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()
    assert np.abs(A @ np.ones(lsize)).max() <= 1e-13 * np.abs(A).max()
