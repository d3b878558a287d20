"""Element restriction in numpy: the checker of bkx_restrict_gather_f64 /
bkx_restrict_scatter_f64.  Test infrastructure only.  The scatter
accumulates in ascending E-vector order (numpy's add.at applies the updates
in index order), the order the kernel sums each node's contributions in."""
import numpy as np


def gather(offsets, uL):
    return np.asarray(uL)[np.asarray(offsets).reshape(-1)]


def scatter(offsets, vE, lsize):
    out = np.zeros(lsize)
    np.add.at(out, np.asarray(offsets).reshape(-1), np.asarray(vE).reshape(-1))
    return out


def assemble(offsets, element_matrices, lsize):
    """Dense assembled matrix sum_e R_e^T A_e R_e."""
    A = np.zeros((lsize, lsize))
    for off, Ae in zip(np.asarray(offsets), element_matrices):
        A[np.ix_(off, off)] += Ae
    return A
