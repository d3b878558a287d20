"""Geometric factors from element nodal coordinates, in numpy with extended
precision (long double): the checker of the geometric-factor kernels
(omp_bk_amd/csrc/geo.hip).  Test infrastructure only.

coords (E, 3, n, n, n) at the GLL nodes; J(c, d) = d x_c / d xi_d of the
degree-p interpolant.  kind 1/3: at the q = p + 2 Gauss points, interpolating
with B(i, p) = l_i(g_p) and differentiating the interpolant with the Gauss
collocation derivative D(n, p) = L_n'(g_p) (exact: the interpolant has degree
p < q); kind 5: at the GLL nodes with D(i, n) = l_n'(x_i).
G = w detJ J^-1 J^-T = w adj(J) adj(J)^T / detJ (rr rs rt ss st tt) and
JxW = w detJ.  The given FP64 coordinates, bases and weights are taken as
exact and everything after them is computed in long double (64-bit
significand on x86), so that the result is the factors of the FP64 inputs
to well below FP64 rounding; it is returned in FP64.  The coordinates are
differentiated relative to the element's first node (subtracted exactly in
long double): the interpolation and derivative matrices annihilate constants
mathematically but not once rounded to FP64, and the kernels do the same."""
import numpy as np

from omp_bk_amd import geometry as geo

SYM = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))
LD = np.longdouble


def jacobian(kind, p, coords):
    """J (E, m, m, m, 3, 3) at the operator's points, in long double."""
    X = np.asarray(coords, dtype=np.float64).astype(LD)
    X = X - X[:, :, :1, :1, :1]
    if kind == 5:
        D = geo.bk5_basis(p).astype(LD)                       # D(i, n)
        d1 = np.einsum("in,ecnjk->ecijk", D, X)
        d2 = np.einsum("jn,ecink->ecijk", D, X)
        d3 = np.einsum("kn,ecijn->ecijk", D, X)
    else:
        B, D = (a.astype(LD) for a in geo.bk13_bases(p))      # B(i, p), D(n, p)
        U = np.einsum("ip,ecijk->ecpjk", B, X)
        U = np.einsum("jq,ecpjk->ecpqk", B, U)
        U = np.einsum("kr,ecpqk->ecpqr", B, U)
        d1 = np.einsum("np,ecnqr->ecpqr", D, U)
        d2 = np.einsum("nq,ecpnr->ecpqr", D, U)
        d3 = np.einsum("nr,ecpqn->ecpqr", D, U)
    J = np.stack([d1, d2, d3], axis=-1)                        # (E, c, a, b, c', d)
    return np.moveaxis(J, 1, -2)                               # (E, a, b, c', c, d)


def _adjugate(J):
    A = np.empty_like(J)
    A[..., 0, 0] = J[..., 1, 1] * J[..., 2, 2] - J[..., 1, 2] * J[..., 2, 1]
    A[..., 0, 1] = J[..., 0, 2] * J[..., 2, 1] - J[..., 0, 1] * J[..., 2, 2]
    A[..., 0, 2] = J[..., 0, 1] * J[..., 1, 2] - J[..., 0, 2] * J[..., 1, 1]
    A[..., 1, 0] = J[..., 1, 2] * J[..., 2, 0] - J[..., 1, 0] * J[..., 2, 2]
    A[..., 1, 1] = J[..., 0, 0] * J[..., 2, 2] - J[..., 0, 2] * J[..., 2, 0]
    A[..., 1, 2] = J[..., 0, 2] * J[..., 1, 0] - J[..., 0, 0] * J[..., 1, 2]
    A[..., 2, 0] = J[..., 1, 0] * J[..., 2, 1] - J[..., 1, 1] * J[..., 2, 0]
    A[..., 2, 1] = J[..., 0, 1] * J[..., 2, 0] - J[..., 0, 0] * J[..., 2, 1]
    A[..., 2, 2] = J[..., 0, 0] * J[..., 1, 1] - J[..., 0, 1] * J[..., 1, 0]
    return A


def factors(kind, p, coords):
    """kind 1: JxW (E, q, q, q); kind 3: G (E, 6, q, q, q); kind 5: G (E, n, n, 6, n)."""
    J = jacobian(kind, p, coords)
    _, w = geo.gll(p + 1) if kind == 5 else geo.gauss(p + 2)
    w = w.astype(LD)
    w3 = w[:, None, None] * w[None, :, None] * w[None, None, :]
    A = _adjugate(J)
    det = J[..., 0, 0] * A[..., 0, 0] + J[..., 0, 1] * A[..., 1, 0] + J[..., 0, 2] * A[..., 2, 0]
    if kind == 1:
        return (w3[None] * det).astype(np.float64)
    s = w3[None] / det
    G = np.stack([s * (A[..., a, 0] * A[..., b, 0] + A[..., a, 1] * A[..., b, 1] + A[..., a, 2] * A[..., b, 2])
                  for a, b in SYM], axis=1).astype(np.float64)
    return G if kind == 3 else np.ascontiguousarray(G.transpose(0, 2, 3, 1, 4))
