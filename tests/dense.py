"""Dense per-element operator matrices built by Kronecker products: the
brute-force restatement the oracle's sum-factorised loops are checked
against.  Test infrastructure only."""
import numpy as np

SYM = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def kron3(a, b, c):
    return np.kron(a, np.kron(b, c))


def bk1_matrix(B, jxw_e):
    """(B x B x B)^T diag(JxW) (B x B x B) for B(i, p) (n x q)."""
    I3 = kron3(B.T, B.T, B.T)                # (q^3, n^3)
    return I3.T @ (jxw_e.reshape(-1)[:, None] * I3)


def _sym_sum(grads, G6):
    """sum_ab grads[a]^T diag(G_ab) grads[b] with G6 = (6, npts) factors."""
    m = grads[0].shape[1]
    A = np.zeros((m, m))
    for f, (a, b) in enumerate(SYM):
        A += grads[a].T @ (G6[f][:, None] * grads[b])
        if a != b:
            A += grads[b].T @ (G6[f][:, None] * grads[a])
    return A


def bk3_matrix(B, D, G_e):
    """BK3 element matrix; D(n, p) (q x q), G_e (6, q, q, q) factor-major."""
    q = D.shape[0]
    Iq = np.eye(q)
    I3 = kron3(B.T, B.T, B.T)
    Dt = D.T                                  # Dt[p, n] = D(n, p)
    grads = [kron3(Dt, Iq, Iq) @ I3, kron3(Iq, Dt, Iq) @ I3, kron3(Iq, Iq, Dt) @ I3]
    return _sym_sum(grads, G_e.reshape(6, -1))


def bk5_matrix(D, G_e):
    """BK5 element matrix; D(i, n) (n x n), G_e (n, n, 6, n) interleaved."""
    n = D.shape[0]
    In = np.eye(n)
    grads = [kron3(D, In, In), kron3(In, D, In), kron3(In, In, D)]
    G6 = np.moveaxis(G_e, 2, 0).reshape(6, -1)
    return _sym_sum(grads, G6)
