"""Bases, meshes and geometric factors for the bake-off operators, in FP64.

The reference runs its operators on synthetic data only: a cos "test basis"
and constant or LCG-filled inputs (include/bk_common.h:60-107).  This module
provides those generators and, for the workloads with real geometry, the
standard spectral-element ingredients:

* Gauss-Legendre points and weights (numpy's ``leggauss``) and
  Gauss-Lobatto-Legendre (GLL) points (the ends and the roots of P'_N) with
  weights 2 / (N (N+1) P_N(x)^2);
* Lagrange interpolation and derivative matrices on any node set;
* the operators' 1D matrices in the reference's index conventions:
    BK1/BK3  B(i, p) = l_i(g_p)   (n GLL modes at q Gauss points; basis[i*q+p])
    BK3      D(n, p) = L_n'(g_p)  (Lagrange on the q Gauss points; dbasis[n*q+p])
    BK5      D(i, n) = l_n'(x_i)  (collocated on the n GLL points; dbasis[i*n+n'])
* a box mesh of the unit cube displaced by
  delta(x) = a h sin(2 pi x) sin(2 pi y) sin(2 pi z) in each of the three
  components (the same displacement for x, y and z; it vanishes on the cube's
  boundary, so the mesh still fills the unit cube); each element map is its
  affine box map followed by x -> x + delta(x) (1, 1, 1);
* the geometric factors at the operator's points, from the exact Jacobian
  J = dx/dxi of that map:  JxW = w detJ (BK1) and
  G = w detJ J^-1 J^-T with factors rr rs rt ss st tt (BK3, BK5).

Element e of an nx x ny x nz box is (ex, ey, ez) with e = (ez ny + ey) nx + ex;
the reference directions (r, s, t) of the tensor indices (i, j, k) map to the
physical axes (x, y, z).
"""
import numpy as np
from numpy.polynomial import legendre as _leg

FACTORS = ("rr", "rs", "rt", "ss", "st", "tt")
_SYM = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


# ---------------------------------------------------------------------------
# the reference's synthetic data (include/bk_common.h)
# ---------------------------------------------------------------------------

def test_basis(rows, cols):
    """M(a, b) = cos(a * cols + b) in double (bk_common.h:63-86, T = double)."""
    a = np.arange(rows * cols, dtype=np.float64)
    return np.cos(a).reshape(rows, cols)


def _lcg_states(n, seed):
    """States s_1..s_n of s <- 1664525 s + 1013904223 (mod 2^32) from s_0 = seed."""
    a, c, m = 1664525, 1013904223, 1 << 32
    out = np.empty(n, dtype=np.uint64)
    blk = min(n, 1 << 16)
    s = seed
    for x in range(blk):
        s = (a * s + c) % m
        out[x] = s
    # jump of blk steps: f^blk(s) = A s + C (mod 2^32)
    A, C = 1, 0
    for _ in range(blk):
        A, C = (a * A) % m, (a * C + c) % m
    A64, C64, mask = np.uint64(A), np.uint64(C), np.uint64(m - 1)
    done = blk
    while done < n:
        take = min(blk, n - done)
        out[done:done + take] = (out[done - blk:done - blk + take] * A64 + C64) & mask
        done += take
    return out


def lcg_fill(n_in, n_jxw, seed=12345):
    """The LCG test data of bk_common.h:92-107: `in` in [-1, 1), then JxW/G = 1.5 + [-1, 1)."""
    s = _lcg_states(n_in + n_jxw, seed)
    v = (s >> np.uint64(8)).astype(np.float64) / float(1 << 23) - 1.0
    return v[:n_in].copy(), 1.5 + v[n_in:]


# ---------------------------------------------------------------------------
# 1D points, weights and Lagrange matrices
# ---------------------------------------------------------------------------

def gauss(q):
    """Gauss-Legendre points and weights on [-1, 1]."""
    x, w = _leg.leggauss(q)
    return np.asarray(x, dtype=np.float64), np.asarray(w, dtype=np.float64)


def gll(n):
    """Gauss-Lobatto-Legendre points (ends and roots of P'_{n-1}) and weights."""
    if n < 2:
        raise ValueError("GLL needs n >= 2 points")
    N = n - 1
    c = np.zeros(N + 1)
    c[N] = 1.0
    dc = _leg.legder(c)
    inner = np.sort(np.real(_leg.legroots(dc))) if N >= 2 else np.zeros(0)
    # two Newton steps on P'_N polish the companion-matrix roots
    ddc = _leg.legder(dc)
    for _ in range(2):
        if inner.size:
            inner = inner - _leg.legval(inner, dc) / _leg.legval(inner, ddc)
    x = np.concatenate(([-1.0], inner, [1.0]))
    w = 2.0 / (N * (N + 1) * _leg.legval(x, c) ** 2)
    return x, w


def lagrange(nodes, x):
    """L[i, p] = l_i(x_p) for the Lagrange polynomials l_i on ``nodes``."""
    nodes = np.asarray(nodes, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    n = nodes.size
    L = np.ones((n, x.size))
    for i in range(n):
        for m in range(n):
            if m != i:
                L[i] *= (x - nodes[m]) / (nodes[i] - nodes[m])
    return L


def lagrange_deriv(nodes, x):
    """dL[i, p] = l_i'(x_p)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    n = nodes.size
    dL = np.zeros((n, x.size))
    for i in range(n):
        for m in range(n):
            if m == i:
                continue
            term = np.full(x.size, 1.0 / (nodes[i] - nodes[m]))
            for l in range(n):
                if l != i and l != m:
                    term *= (x - nodes[l]) / (nodes[i] - nodes[l])
            dL[i] += term
    return dL


def bk13_bases(p):
    """(B, D) of BK1/BK3 at order p: B (n x q) = l_i^GLL(g_p), D (q x q) = L_n^Gauss'(g_p)."""
    n, q = p + 1, p + 2
    xg, _ = gauss(q)
    xl, _ = gll(n)
    return lagrange(xl, xg), lagrange_deriv(xg, xg)


def bk5_basis(p):
    """D (n x n) of BK5 at order p: D(i, n') = l_n'(x_i) on the n = p + 1 GLL points."""
    xl, _ = gll(p + 1)
    return lagrange_deriv(xl, xl).T.copy()


# ---------------------------------------------------------------------------
# meshes and geometric factors
# ---------------------------------------------------------------------------

def box_shape(nelmt):
    """An nx x ny x nz box with nx >= ny >= nz, nx ny nz = nelmt (a power of two)."""
    k = int(round(np.log2(nelmt)))
    if 1 << k != nelmt:
        raise ValueError("box meshes need a power-of-two element count")
    nz = 1 << (k // 3)
    ny = 1 << ((k - k // 3) // 2)
    nx = nelmt // (ny * nz)
    return nx, ny, nz


def _axes(shape, pts, e0, e1, xp, dev):
    """Undisplaced coordinates of the points pts (reference coordinates on
    [-1, 1]) of elements e0..e1-1, one array per axis, shaped to broadcast to
    (e1 - e0, m, m, m): x (., m, 1, 1), y (., 1, m, 1), z (., 1, 1, m).
    xp is numpy or torch (then on device dev)."""
    nx, ny, nz = shape
    if xp is np:
        t = (np.asarray(pts, dtype=np.float64) + 1.0) / 2.0
        elems = np.arange(e0, e1, dtype=np.int64)
    else:
        t = (xp.as_tensor(pts, dtype=xp.float64, device=dev) + 1.0) / 2.0
        elems = xp.arange(e0, e1, dtype=xp.int64, device=dev)
    ex = elems % nx
    ey = (elems // nx) % ny
    ez = elems // (nx * ny)
    X = (ex[:, None] + t[None, :]) / nx
    Y = (ey[:, None] + t[None, :]) / ny
    Z = (ez[:, None] + t[None, :]) / nz
    return X[:, :, None, None], Y[:, None, :, None], Z[:, None, None, :]


def displacement(x, y, z, amp, h, xp=np):
    """delta(x) = amp h sin(2 pi x) sin(2 pi y) sin(2 pi z) and its gradient
    (x, y, z broadcast against each other)."""
    k = 2 * np.pi
    sx, sy, sz = xp.sin(k * x), xp.sin(k * y), xp.sin(k * z)
    cx, cy, cz = xp.cos(k * x), xp.cos(k * y), xp.cos(k * z)
    a = amp * h
    return a * sx * sy * sz, (a * k * cx * sy * sz, a * k * sx * cy * sz, a * k * sx * sy * cz)


def _factors(shape, pts, w, e0, e1, amp, xp, dev):
    """(w detJ, G = w detJ J^-1 J^-T as 6 factors (e1-e0, 6, m, m, m)) of the map
    phi(x) = x + delta(x) (1, 1, 1) composed with the element's affine map.

    J = (I + 1 g^T) H with g = grad delta and H = diag(hx, hy, hz) / 2, so
    det J = det H (1 + g.1) (matrix determinant lemma) and
    J^-1 = H^-1 (I - 1 g^T / (1 + g.1)) (Sherman-Morrison), hence
    J^-1 J^-T = H^-1 (I - (1 g^T + g 1^T) / s + (g.g) 1 1^T / s^2) H^-1."""
    nx, ny, nz = shape
    hh = (0.5 / nx, 0.5 / ny, 0.5 / nz)
    X, Y, Z = _axes(shape, pts, e0, e1, xp, dev)
    _, g = displacement(X, Y, Z, amp, 1.0 / max(shape), xp)
    s = 1.0 + g[0] + g[1] + g[2]
    w3 = w[:, None, None] * w[None, :, None] * w[None, None, :]
    if xp is not np:
        w3 = xp.as_tensor(w3, dtype=xp.float64, device=dev)
    wd = w3[None] * (hh[0] * hh[1] * hh[2]) * s
    rs = 1.0 / s
    gs = [gi * rs for gi in g]
    gg = gs[0] * gs[0] + gs[1] * gs[1] + gs[2] * gs[2]
    G = []
    for a, b in _SYM:
        kk = gg - gs[a] - gs[b]
        if a == b:
            kk = kk + 1.0
        G.append(wd * kk * (1.0 / (hh[a] * hh[b])))
    return wd, xp.stack(G, 1)


def geometric_factors(kind, p, nelmt, amp=0.1, shape=None, chunk=8192, coords=False, device=None):
    """Geometric factors of a perturbed box mesh in the layout of operator ``kind``.

    kind 1: JxW (E, q, q, q);  kind 3: G (E, 6, q, q, q) factor-major;
    kind 5: G (E, n, n, 6, n) interleaved (GLL-collocated points).
    The element map is the affine box map followed by the smooth displacement
    x -> x + delta(x) (1, 1, 1), with its exact Jacobian.  With ``coords``
    also returns the physical coordinates (E, m, m, m, 3) of the points
    (Gauss points for 1/3, GLL nodes for 5).  With ``device`` the arrays are
    float64 torch tensors computed on that device, else numpy arrays."""
    shape = shape or box_shape(nelmt)
    if shape[0] * shape[1] * shape[2] != nelmt:
        raise ValueError("mesh shape does not hold nelmt elements")
    if kind not in (1, 3, 5):
        raise ValueError("kind must be 1, 3 or 5")
    pts, w = gll(p + 1) if kind == 5 else gauss(p + 2)
    m = pts.size
    dims = {1: (nelmt, m, m, m), 3: (nelmt, 6, m, m, m), 5: (nelmt, m, m, 6, m)}[kind]
    if device is None:
        xp, dev = np, None
        out = np.empty(dims)
        xyz = np.empty((nelmt, m, m, m, 3)) if coords else None
    else:
        import torch as xp
        dev = device
        out = xp.empty(dims, dtype=xp.float64, device=dev)
        xyz = xp.empty((nelmt, m, m, m, 3), dtype=xp.float64, device=dev) if coords else None
        chunk = max(chunk, 65536)
    h = 1.0 / max(shape)
    for e0 in range(0, nelmt, chunk):
        e1 = min(nelmt, e0 + chunk)
        wd, G = _factors(shape, pts, w, e0, e1, amp, xp, dev)
        if kind == 1:
            out[e0:e1] = wd
        elif kind == 3:
            out[e0:e1] = G
        else:
            out[e0:e1] = G.permute(0, 2, 3, 1, 4) if xp is not np else G.transpose(0, 2, 3, 1, 4)
        if coords:
            X, Y, Z = _axes(shape, pts, e0, e1, xp, dev)
            d, _ = displacement(X, Y, Z, amp, h, xp)
            for c, axis in enumerate((X, Y, Z)):
                xyz[e0:e1, ..., c] = axis + d
    return (out, xyz) if coords else out


def nodal_coords(p, nelmt, amp=0.1, shape=None, device=None):
    """Coordinates (E, 3, n, n, n) of the perturbed box mesh at each element's
    n = p + 1 GLL nodes (the input of the geometric-factor kernels): the
    affine box map followed by x -> x + delta(x) (1, 1, 1).  numpy, or a
    float64 torch tensor on ``device``."""
    shape = shape or box_shape(nelmt)
    if shape[0] * shape[1] * shape[2] != nelmt:
        raise ValueError("mesh shape does not hold nelmt elements")
    pts, _ = gll(p + 1)
    if device is None:
        xp, dev = np, None
    else:
        import torch as xp
        dev = device
    X, Y, Z = _axes(shape, pts, 0, nelmt, xp, dev)
    d, _ = displacement(X, Y, Z, amp, 1.0 / max(shape), xp)
    comps = [X + d, Y + d, Z + d]
    if xp is np:
        return np.stack([np.broadcast_to(c, d.shape) for c in comps], axis=1).copy()
    return xp.stack([c.expand(d.shape) for c in comps], dim=1).contiguous()


def box_offsets(p, shape):
    """Element-restriction offsets of the continuous degree-p space on an
    nx x ny x nz box mesh: int32 (E, n^3), the global node of each element's
    GLL node (i, j, k), numbered over the (nx p + 1) x (ny p + 1) x (nz p + 1)
    grid with z fastest (the axis of the element's fastest index k, so that
    consecutive global nodes are consecutive E-vector values within an
    element and the gather and scatter read whole lines); and the number of
    global nodes."""
    nx, ny, nz = shape
    n = p + 1
    NX, NY, NZ = nx * p + 1, ny * p + 1, nz * p + 1
    e = np.arange(nx * ny * nz)
    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
    i = np.arange(n)
    gx = (ex[:, None] * p + i[None, :])[:, :, None, None]
    gy = (ey[:, None] * p + i[None, :])[:, None, :, None]
    gz = (ez[:, None] * p + i[None, :])[:, None, None, :]
    off = (gx * NY + gy) * NZ + gz
    return off.reshape(nx * ny * nz, n ** 3).astype(np.int32), NX * NY * NZ


def random_u(nvalues, seed=0):
    """i.i.d. uniform[-1, 1] FP64 values from numpy's PCG64 with ``seed``."""
    return np.random.default_rng(seed).uniform(-1.0, 1.0, nvalues)


def smooth_mode(xyz):
    """sin(pi x) sin(pi y) sin(pi z) at the points xyz[..., 3] (numpy or torch)."""
    sin = np.sin if isinstance(xyz, np.ndarray) else type(xyz).sin
    return sin(np.pi * xyz[..., 0]) * sin(np.pi * xyz[..., 1]) * sin(np.pi * xyz[..., 2])
