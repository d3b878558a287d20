"""Mixed-order BK3 batches: workload, plan, launches and tuning.

A batch holds elements of orders p = 1..8 in one E-vector (layout in
include/bkx.h: element e's values at uoff[e], its factors at goff[e], the
bases of all orders packed).  The workload of the benchmark: 200,000
elements whose orders follow a Zipf(s = 1.2) distribution over a seeded
shuffled list of the eight orders, each element with its own
perturbed-affine geometry (x = x0 + A_e xi, A_e = h (I + a R_e), R_e uniform
in [-1, 1]^{3x3}) and FP64 uniform[-1, 1] inputs.

The batch runs as one launch over all orders or as one launch per order
(each order by a kernel compiled for it alone), either captured in a
hipGraph.  The elements per workgroup (divisor 1, 2 or 4 of
floor(256 / q^2)) are tuned per order for each: ``tune_single`` for the
single launch, ``tune`` for the per-order launches (with the rows of G in
flight).
"""
import ctypes
import statistics

import numpy as np
import torch

from . import _lib
from . import geometry as geo

_vp = ctypes.c_void_p
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def _bind():
    L = _lib.lib()
    if getattr(L, "_bkx_mixed_bound", False):
        return L
    L.bkx_mixed_plan_create.argtypes = [_i32p, ctypes.c_size_t, _i32p, _i32p, ctypes.c_int, ctypes.POINTER(_vp)]
    L.bkx_mixed_plan_create_grouped.argtypes = L.bkx_mixed_plan_create.argtypes
    L.bkx_mixed_plan_destroy.argtypes = [_vp]
    L.bkx_mixed_plan_info.argtypes = [_vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.bkx_mixed_plan_offsets.argtypes = [_vp, _i64p, _i64p]
    L.bkx_mixed_plan_set_order_shape.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.bkx_mixed_plan_get_order_shape.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                                 ctypes.POINTER(ctypes.c_int)]
    L.bkx_bk3_mixed_apply_f64.argtypes = [_vp] + [_vp] * 5
    L.bkx_bk3_mixed_apply_by_order_f64.argtypes = [_vp] + [_vp] * 5
    L.bkx_bk3_mixed_apply_order_f64.argtypes = [_vp, ctypes.c_int] + [_vp] * 5
    L.bkx_bk3_mixed_graph_create.argtypes = [_vp] + [_vp] * 4 + [ctypes.c_int, ctypes.POINTER(_vp)]
    L.bkx_bk3_mixed_prepared_size.argtypes = []
    L.bkx_bk3_mixed_prepared_size.restype = ctypes.c_size_t
    L.bkx_bk3_mixed_prepare.argtypes = [_vp, _vp]
    L.bkx_bk3_mixed_apply_prepared_f64.argtypes = [_vp] + [_vp] * 5
    L.bkx_bk3_mixed_apply_by_order_prepared_f64.argtypes = [_vp] + [_vp] * 5
    L.bkx_bk3_mixed_graph_create_prepared.argtypes = [_vp] + [_vp] * 4 + [ctypes.c_int, ctypes.POINTER(_vp)]
    L.bkx_graph_launch.argtypes = [_vp, _vp]
    L.bkx_graph_destroy.argtypes = [_vp]
    for name in ("bkx_mixed_plan_create", "bkx_mixed_plan_create_grouped", "bkx_mixed_plan_destroy", "bkx_mixed_plan_info", "bkx_mixed_plan_offsets",
                 "bkx_mixed_plan_set_order_shape", "bkx_mixed_plan_get_order_shape",
                 "bkx_bk3_mixed_apply_f64", "bkx_bk3_mixed_apply_by_order_f64", "bkx_bk3_mixed_apply_order_f64",
                 "bkx_bk3_mixed_graph_create", "bkx_graph_launch", "bkx_graph_destroy", "bkx_bk3_mixed_prepare",
                 "bkx_bk3_mixed_apply_prepared_f64", "bkx_bk3_mixed_apply_by_order_prepared_f64",
                 "bkx_bk3_mixed_graph_create_prepared"):
        getattr(L, name).restype = ctypes.c_int
    L._bkx_mixed_bound = True
    return L


def _check(name, status):
    if status != 0:
        raise _lib.BkxError(name, status, _lib.lib().bkx_status_string(status).decode())


# ---------------------------------------------------------------------------
# workload
# ---------------------------------------------------------------------------

def zipf_orders(nelmt=200000, s=1.2, seed=2024):
    """Orders of nelmt elements: rank k = 1..8 of a seeded shuffled list of
    the orders 1..8 gets the Zipf(s) share k^-s / sum_j j^-s of the elements
    (rounded, the remainder to rank 1), and the element order is shuffled."""
    rng = np.random.default_rng(seed)
    ranked = rng.permutation(np.arange(1, 9))
    w = np.arange(1, 9, dtype=np.float64) ** -s
    w /= w.sum()
    counts = np.floor(w * nelmt).astype(np.int64)
    counts[0] += nelmt - counts.sum()
    order = np.repeat(ranked, counts).astype(np.int32)
    rng.shuffle(order)
    return order


# doubles of the packed bases: B_p (n x q) and D_p (q x q) for p = 1..8
PACKED_BASES = sum((p + 1) * (p + 2) + (p + 2) ** 2 for p in range(1, 9))


def prepared_size():
    """Doubles of the even-odd forms of the packed bases."""
    return int(_bind().bkx_bk3_mixed_prepared_size())


def packed_bases():
    """B_p (n x q) then D_p (q x q) for p = 1..8 (the layout of include/bkx.h)."""
    parts = []
    for p in range(1, 9):
        B, D = geo.bk13_bases(p)
        parts += [B.reshape(-1), D.reshape(-1)]
    return np.concatenate(parts)


def prepared_bases(bases):
    """The even-odd forms of the packed bases of every order
    (bkx_bk3_mixed_prepare, on the host), on the device of ``bases``; raises
    ValueError when an order's bases are not centro-symmetric."""
    L = _bind()
    host = np.ascontiguousarray(torch.as_tensor(bases).detach().cpu().numpy(), dtype=np.float64).reshape(-1)
    out = np.empty(L.bkx_bk3_mixed_prepared_size())
    st = L.bkx_bk3_mixed_prepare(host.ctypes.data, out.ctypes.data)
    if st == 5:
        raise ValueError("the bases are not centro-symmetric")
    _check("bkx_bk3_mixed_prepare", st)
    device = bases.device if isinstance(bases, torch.Tensor) else None
    return torch.as_tensor(out, device=device)


def affine_factors(order, goff, ngeo, amp=0.1, seed=7, device=None):
    """G of every element of a mixed batch: element e is x = x0 + A_e xi with
    A_e = h (I + amp R_e), h = 1 / cbrt(nelmt); G = w detJ J^-1 J^-T (J = A_e,
    constant over the element) at its Gauss points, factor-major.  With
    ``device`` the factors are scattered into a device tensor there (the
    per-element matrices are made on the host either way)."""
    nelmt = order.size
    rng = np.random.default_rng(seed)
    h = 1.0 / np.cbrt(nelmt)
    A = h * (np.eye(3)[None] + amp * rng.uniform(-1, 1, (nelmt, 3, 3)))
    inv = np.linalg.inv(A)
    M = np.einsum("eab,ecb->eac", inv, inv) * np.linalg.det(A)[:, None, None]
    sym = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    M6 = np.stack([M[:, a, b] for a, b in sym], axis=1)      # (nelmt, 6)
    if device is None:
        G = np.empty(ngeo)
    else:
        G = torch.empty(ngeo, dtype=torch.float64, device=device)
        M6_d = torch.as_tensor(M6, device=device)
        goff_d = torch.as_tensor(goff, device=device)
    for p in range(1, 9):
        idx = np.nonzero(order == p)[0]
        if idx.size == 0:
            continue
        q = p + 2
        _, w = geo.gauss(q)
        w3 = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
        if device is None:
            vals = M6[idx][:, :, None] * w3[None, None, :]              # (k, 6, q^3)
            pos = goff[idx][:, None] + np.arange(6 * q ** 3)[None, :]
            G[pos] = vals.reshape(idx.size, -1)
        else:
            ii = torch.as_tensor(idx, device=device)
            vals = M6_d[ii][:, :, None] * torch.as_tensor(w3, device=device)[None, None, :]
            pos = goff_d[ii][:, None] + torch.arange(6 * q ** 3, device=device)[None, :]
            G[pos.reshape(-1)] = vals.reshape(-1)
    return G


# ---------------------------------------------------------------------------
# plan
# ---------------------------------------------------------------------------

class MixedPlan:
    """The device-side tile lists of one batch (bkx_mixed_plan_create).  With
    grouped, the batch's arrays are in the grouped layout
    (bkx_mixed_plan_create_grouped): each order's elements contiguous, each
    order's range on a 256-byte boundary; uoff and goff give every element's
    place either way.  With rounds_by_order or a grid the single launch is
    persistent: its workgroups take tiles from a queue held by the plan, so
    launches of one such plan (direct or replayed) must not overlap."""

    def __init__(self, order, div_by_order=None, rounds_by_order=None, grid=0, grouped=False):
        L = _bind()
        self.order = np.ascontiguousarray(order, dtype=np.int32)
        self.nelmt = self.order.size
        div = None if div_by_order is None else np.ascontiguousarray(div_by_order, dtype=np.int32)
        rnd = None if rounds_by_order is None else np.ascontiguousarray(rounds_by_order, dtype=np.int32)
        for name, a in (("div_by_order", div), ("rounds_by_order", rnd)):
            if a is not None and a.shape != (9,):
                raise ValueError("%s is indexed by the order: 9 entries, [0] unused and [p] for p = 1..8" % name)
        self._div, self._rnd = div, rnd
        h = _vp()
        name = "bkx_mixed_plan_create_grouped" if grouped else "bkx_mixed_plan_create"
        _check(name, getattr(L, name)(
            self.order.ctypes.data_as(_i32p), self.nelmt, None if div is None else div.ctypes.data_as(_i32p),
            None if rnd is None else rnd.ctypes.data_as(_i32p), int(grid), ctypes.byref(h)))
        self.handle = h
        self.grouped = bool(grouped)
        self.dofs = int(np.sum((self.order.astype(np.int64) + 1) ** 3))   # without padding
        nd, ng, nt, gr = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int(), ctypes.c_int()
        _check("bkx_mixed_plan_info", L.bkx_mixed_plan_info(h, ctypes.byref(nd), ctypes.byref(ng), ctypes.byref(nt),
                                                            ctypes.byref(gr)))
        self.ndof, self.ngeo, self.ntiles, self.grid = nd.value, ng.value, nt.value, gr.value
        self.uoff = np.empty(self.nelmt, dtype=np.int64)
        self.goff = np.empty(self.nelmt, dtype=np.int64)
        _check("bkx_mixed_plan_offsets", L.bkx_mixed_plan_offsets(h, self.uoff.ctypes.data_as(_i64p),
                                                                  self.goff.ctypes.data_as(_i64p)))

    def set_order_shape(self, p, div, gpf):
        """Launch shape of order p's kernel in the per-order launches:
        elements-per-workgroup divisor div (1, 2, 4) and rows of G in flight
        gpf (2, 3, 4)."""
        _check("bkx_mixed_plan_set_order_shape", _bind().bkx_mixed_plan_set_order_shape(self.handle, p, div, gpf))

    def order_shape(self, p):
        d, g = ctypes.c_int(), ctypes.c_int()
        _check("bkx_mixed_plan_get_order_shape",
               _bind().bkx_mixed_plan_get_order_shape(self.handle, p, ctypes.byref(d), ctypes.byref(g)))
        return d.value, g.value

    def close(self):
        if self.handle:
            _bind().bkx_mixed_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _args(self, bases, G, u, v, prepared=False):
        # bases: the packed bases, or with prepared their even-odd forms
        # (prepared_bases); the kernels read every order's entries, so a
        # short array would be read out of bounds on the device
        for name, t in (("bases", bases), ("G", G), ("u", u), ("v", v)):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
                raise TypeError("%s must be a contiguous float64 tensor on a HIP device" % name)
        if u.numel() != self.ndof or v.numel() != self.ndof or G.numel() != self.ngeo:
            raise ValueError("u and v must hold %d values and G %d" % (self.ndof, self.ngeo))
        nb = prepared_size() if prepared else PACKED_BASES
        if bases.numel() != nb:
            raise ValueError("bases must hold %d doubles (%s), got %d"
                             % (nb, "even-odd forms" if prepared else "packed B and D of p = 1..8", bases.numel()))
        return bases.data_ptr(), G.data_ptr(), u.data_ptr(), v.data_ptr()

    def apply(self, bases, G, u, v, prepared=False):
        """One launch over all tiles, on the current torch stream; with
        prepared, bases are the even-odd forms of prepared_bases."""
        name = "bkx_bk3_mixed_apply_prepared_f64" if prepared else "bkx_bk3_mixed_apply_f64"
        _check(name, getattr(_bind(), name)(self.handle, *self._args(bases, G, u, v, prepared),
                                            torch.cuda.current_stream().cuda_stream))
        return v

    def apply_by_order(self, bases, G, u, v, prepared=False):
        """One launch per order."""
        name = "bkx_bk3_mixed_apply_by_order_prepared_f64" if prepared else "bkx_bk3_mixed_apply_by_order_f64"
        _check(name, getattr(_bind(), name)(self.handle, *self._args(bases, G, u, v, prepared),
                                            torch.cuda.current_stream().cuda_stream))
        return v

    def apply_order(self, p, bases, G, u, v):
        _check("bkx_bk3_mixed_apply_order_f64", _bind().bkx_bk3_mixed_apply_order_f64(
            self.handle, p, *self._args(bases, G, u, v), torch.cuda.current_stream().cuda_stream))
        return v

    def graph(self, bases, G, u, v, by_order=False, prepared=False):
        """A hipGraph of the launch(es) for these arrays (MixedGraph)."""
        return MixedGraph(self, bases, G, u, v, by_order, prepared)


class MixedGraph:
    def __init__(self, plan, bases, G, u, v, by_order, prepared=False):
        h = _vp()
        name = "bkx_bk3_mixed_graph_create_prepared" if prepared else "bkx_bk3_mixed_graph_create"
        _check(name, getattr(_bind(), name)(plan.handle, *plan._args(bases, G, u, v, prepared), int(by_order),
                                            ctypes.byref(h)))
        self.handle = h
        self._keep = (plan, bases, G, u, v)

    def launch(self):
        _check("bkx_graph_launch", _bind().bkx_graph_launch(self.handle, torch.cuda.current_stream().cuda_stream))

    def set_order_shape(self, p, div, gpf):
        """Launch shape of order p's kernel in the per-order launches:
        elements-per-workgroup divisor div (1, 2, 4) and rows of G in flight
        gpf (2, 3, 4)."""
        _check("bkx_mixed_plan_set_order_shape", _bind().bkx_mixed_plan_set_order_shape(self.handle, p, div, gpf))

    def order_shape(self, p):
        d, g = ctypes.c_int(), ctypes.c_int()
        _check("bkx_mixed_plan_get_order_shape",
               _bind().bkx_mixed_plan_get_order_shape(self.handle, p, ctypes.byref(d), ctypes.byref(g)))
        return d.value, g.value

    def close(self):
        if self.handle:
            _bind().bkx_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _median_ms(fn, steps=10, warmup=2):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    out = []
    for _ in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        fn()
        b.record(s)
        b.synchronize()
        out.append(a.elapsed_time(b))
    return statistics.median(out)


def _batch_ms(fn, steps=10, reps=10, warmup=2):
    """Median over steps of the mean time of reps back-to-back calls: the
    steady state of a launch inside a sequence of launches, which for short
    launches differs from a launch timed alone on an idle GPU."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    out = []
    for _ in range(steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        for _ in range(reps):
            fn()
        b.record(s)
        b.synchronize()
        out.append(a.elapsed_time(b) / reps)
    return statistics.median(out)


ORDER_SHAPES = [(d, g) for d in (1, 2, 4) for g in (2, 3, 4)]


def tune(order, bases, G, u, v, steps=10, plan=None):
    """For the per-order launches: per order, the launch shape (elements-
    per-workgroup divisor 1, 2 or 4, and rows of G in flight 2, 3 or 4)
    whose launch of that order's elements (by the kernel compiled for that
    order) is fastest, timed back to back as in the per-order sequence.
    With plan (the plan whose layout u, v and G follow), the shapes are
    timed on it and set on it; else on a plan of the caller's layout.
    Returns ({p: (div, gpf)}, {p: {(div, gpf): ms}})."""
    own = plan is None
    pl = MixedPlan(order) if own else plan
    best, times = {}, {}
    for p in range(1, 9):
        if not np.any(order == p):
            continue
        times[p] = {}
        for div, gpf in ORDER_SHAPES:
            pl.set_order_shape(p, div, gpf)
            times[p][(div, gpf)] = _batch_ms(lambda: pl.apply_order(p, bases, G, u, v), steps)
        best[p] = min(times[p], key=times[p].get)
        pl.set_order_shape(p, *best[p])
    if own:
        pl.close()
    return best, times


def tune_single(order, bases, G, u, v, steps=10):
    """For the single launch over all orders, whose kernel is sized for its
    largest order body: the divisors per order, chosen greedily over the
    orders present, most elements first, each set to 2 or 4 when the whole
    launch gets faster; returns (div_by_order, ms of the chosen plan)."""
    counts = np.bincount(np.asarray(order), minlength=9)

    def timed(div):
        pl = MixedPlan(order, div_by_order=div)
        try:
            return _median_ms(lambda: pl.apply(bases, G, u, v), steps)
        finally:
            pl.close()

    best = [1] * 9
    best_ms = timed(best)
    for p in sorted((p for p in range(1, 9) if counts[p] > 0), key=lambda p: -counts[p]):
        base = list(best)
        for div in (2, 4):
            trial = list(base)
            trial[p] = div
            ms = timed(trial)
            if ms < best_ms:
                best, best_ms = trial, ms
    return best, best_ms
