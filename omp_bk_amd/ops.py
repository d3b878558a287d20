"""Host-side mirror of the reference's operator interface on torch tensors.

The reference's kernels have the signatures
  BK1  SumFactorization<T, nq>(nelmt, basis, JxW, in, out)       openmp/bk1_sumfact.h:16-22
  BK3  SumFactorization<T, nq>(nelmt, basis, dbasis, G, in, out) openmp/bk3_sumfact.h:17-24
  BK5  SumFactorization<T, nq>(nelmt, dbasis, G, in, out)        openmp/BK5.cpp:18-24
with the order fixed at compile time.  Here the order p is an argument,
the arrays are contiguous float64 tensors on one HIP device (flat or shaped,
as long as their sizes match the element count), and the kernels run on the
current torch stream through the C ABI of include/bkx.h.  Like the reference
the calls return once the work is enqueued on the stream; torch's stream
semantics order later work after it.
"""
import contextlib

import numpy as np
import torch

from . import _lib

# the current stream's raw handle by device index (torch's own accessor,
# about 0.1 us against 2 us for torch.cuda.current_stream(device))
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _check(name, t, numel, dtype=torch.float64):
    # the common case in one expression; the checks below name the failure
    if isinstance(t, torch.Tensor) and t.is_cuda and t.dtype is dtype and t.is_contiguous() and t.numel() == numel:
        return
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch.Tensor" % name)
    if t.device.type != "cuda":
        raise ValueError("%s must be on a HIP device, got %s" % (name, t.device))
    if t.dtype != dtype:
        raise TypeError("%s must be %s, got %s" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if t.numel() != numel:
        raise ValueError("%s has %d values, expected %d" % (name, t.numel(), numel))


def _apart(out, u):
    """Raises on an in-place apply (out at u's address): the kernels read u
    while other workgroups write out.  Only that case is checked, to keep the
    per-call host cost low (0.1 us); any other overlap of an output with an
    input is as undefined as in the C ABI."""
    if out.data_ptr() == u.data_ptr() and u.numel():
        raise ValueError("out must not be u (no in-place apply)")


def _stream(t):
    """The handle of the current torch stream on t's device."""
    idx = t.get_device()
    if _raw_stream is not None:
        return _raw_stream(idx)
    return ctypes_stream(torch.cuda.current_stream(idx))


def _on(t):
    """Makes t's device the current device for a call (nothing to do when
    it is already)."""
    idx = t.get_device()
    if idx == torch.cuda.current_device():
        return contextlib.nullcontext()
    return torch.cuda.device(idx)


def _launch(entry, args, t):
    """Enqueues a checked C-ABI call on the current stream of t's device."""
    with _on(t):
        _lib.call(entry, *args, _stream(t))


class Bound:
    """One operator apply whose arguments were checked once (bind_bk1,
    bind_bk3, bind_bk5).  Each call enqueues the kernel on the current stream
    of the arrays' device with the same arrays, which the object keeps
    alive; their contents may change between calls, their storage must not.
    For loops over one small operator, where the plain wrappers' per-call
    checks take longer on the host than the apply takes on the device.
    Returns the output array."""

    def __init__(self, entry, args, arrays, out):
        self._entry = entry
        self._fn = getattr(_lib.lib(), entry)
        self._args = args
        self._arrays = arrays      # the pointers in _args are theirs
        self._dev = out.get_device()
        self.out = out

    def __call__(self):
        dev = self._dev
        if dev == torch.cuda.current_device():
            status = self._fn(*self._args, _stream(self.out))
        else:
            with torch.cuda.device(dev):
                status = self._fn(*self._args, _stream(self.out))
        if status != 0:
            raise _lib.BkxError(self._entry, status, _lib.lib().bkx_status_string(status).decode())
        return self.out


def ctypes_stream(stream):
    return stream.cuda_stream


_orders = {}


def _order_ok(op, p):
    if op not in _orders:
        _orders[op] = (_lib.lib().bkx_min_order(op), _lib.lib().bkx_max_order(op))
    lo, hi = _orders[op]
    if not (lo <= p <= hi):
        raise ValueError("unsupported polynomial order p = %d (BK%d supports %d..%d)" % (p, op, lo, hi))


# form: (entry-point suffix, storage dtype of u and JxW/G); v is FP64 in all
_BK13_FORMS = {
    "f64": ("f64", torch.float64),
    "ff": ("ff", torch.float64),
    "f32c": ("f32c", torch.float32),
    "f32w": ("f32w", torch.float32),
}

def _bk13_form(op, form):
    if form not in _BK13_FORMS:
        raise ValueError("unknown BK%d form %r (one of %s)" % (op, form, ", ".join(_BK13_FORMS)))
    suffix, sdt = _BK13_FORMS[form]
    return "bkx_bk%d_apply_%s" % (op, suffix), sdt


def bk1(p, basis, JxW, u, out=None, form="f64"):
    """BK1 mass operator v = B^T diag(JxW) B u per element.

    form "f64": u, JxW, v in FP64 (the drop-in);
         "ff": FP64 storage, float-float arithmetic (u and JxW split into
               hi + lo when loaded; the reference's BK1_ff,
               openmp/BK1_ff.cpp:73-723), v FP64;
         "f32c": u, JxW in FP32, float-float compensated arithmetic, v FP64;
         "f32w": u, JxW in FP32, FP64 arithmetic, v FP64."""
    entry, args, out = _bk1_call(p, basis, JxW, u, out, form)
    _launch(entry, args, u)
    return out


def _bk1_call(p, basis, JxW, u, out, form):
    """bk1's checks: its C-ABI entry point, the arguments before the stream,
    and out (allocated when None)."""
    _order_ok(1, p)
    entry, sdt = _bk13_form(1, form)
    nm, nq = p + 1, p + 2
    if u.numel() % nm ** 3:
        raise ValueError("u must hold whole elements of %d values" % nm ** 3)
    ne = u.numel() // nm ** 3
    _check("basis", basis, nm * nq)
    _check("JxW", JxW, ne * nq ** 3, sdt)
    _check("u", u, ne * nm ** 3, sdt)
    if out is None:
        out = torch.empty(u.shape, dtype=torch.float64, device=u.device)
    _check("out", out, ne * nm ** 3)
    _apart(out, u)
    return entry, (p, ne, basis.data_ptr(), JxW.data_ptr(), u.data_ptr(), out.data_ptr()), out


def bk3(p, basis, dbasis, G, u, out=None, form="f64"):
    """BK3 Poisson operator v = B^T D^T G D B u per element.

    form "f64": u, G, v in FP64 (the drop-in);
         "ff": FP64 storage, float-float arithmetic (u and G split into
               hi + lo when loaded), v FP64;
         "f32c": u, G in FP32, float-float compensated arithmetic, v FP64;
         "f32w": u, G in FP32, FP64 arithmetic, v FP64."""
    entry, args, out = _bk3_call(p, basis, dbasis, G, u, out, form)
    _launch(entry, args, u)
    return out


def _bk3_call(p, basis, dbasis, G, u, out, form):
    """bk3's checks, as _bk1_call."""
    _order_ok(3, p)
    entry, sdt = _bk13_form(3, form)
    nm, nq = p + 1, p + 2
    if u.numel() % nm ** 3:
        raise ValueError("u must hold whole elements of %d values" % nm ** 3)
    ne = u.numel() // nm ** 3
    _check("basis", basis, nm * nq)
    _check("dbasis", dbasis, nq * nq)
    _check("G", G, ne * 6 * nq ** 3, sdt)
    _check("u", u, ne * nm ** 3, sdt)
    if out is None:
        out = torch.empty(u.shape, dtype=torch.float64, device=u.device)
    _check("out", out, ne * nm ** 3)
    _apart(out, u)
    return entry, (p, ne, basis.data_ptr(), dbasis.data_ptr(), G.data_ptr(), u.data_ptr(), out.data_ptr()), out


def bk3_basis(p, basis, dbasis, device=None):
    """The even-odd form of centro-symmetric BK3 bases (B(i, p) =
    B(n-1-i, q-1-p), D(m, p) = -D(q-1-m, q-1-p): the GLL-to-Gauss basis and
    the Gauss collocation derivative), built on the host by
    bkx_bk3_basis_prepare, as an FP64 tensor on ``device`` (default: that of
    ``basis`` when it is a tensor).  Raises ValueError when the bases are not
    centro-symmetric."""
    _order_ok(3, p)
    if device is None and isinstance(basis, torch.Tensor):
        device = basis.device
    b = np.ascontiguousarray(torch.as_tensor(basis).detach().cpu().numpy(), dtype=np.float64).reshape(-1)
    d = np.ascontiguousarray(torch.as_tensor(dbasis).detach().cpu().numpy(), dtype=np.float64).reshape(-1)
    nm, nq = p + 1, p + 2
    if b.size != nm * nq or d.size != nq * nq:
        raise ValueError("bases of order %d are %d x %d and %d x %d" % (p, nm, nq, nq, nq))
    out = np.empty(_lib.lib().bkx_bk3_basis_size(p))
    st = _lib.lib().bkx_bk3_basis_prepare(p, b.ctypes.data, d.ctypes.data, out.ctypes.data)
    if st == 5:
        raise ValueError("the bases are not centro-symmetric: use bk3 with basis and dbasis")
    if st != 0:
        raise _lib.BkxError("bkx_bk3_basis_prepare", st, _lib.lib().bkx_status_string(st).decode())
    return torch.as_tensor(out, device=device)


def bk3_prepared(p, prepared, G, u, out=None, form="f64"):
    """BK3 as bk3, with the bases in the even-odd form of bk3_basis; form
    "f64" or "f32w".  The result agrees with bk3's to rounding."""
    _order_ok(3, p)
    entry = {"f64": "bkx_bk3_apply_prepared_f64", "f32w": "bkx_bk3_apply_prepared_f32w"}.get(form)
    if entry is None:
        raise ValueError("form must be f64 or f32w")
    sdt = torch.float64 if form == "f64" else torch.float32
    nm, nq = p + 1, p + 2
    if u.numel() % nm ** 3:
        raise ValueError("u must hold whole elements of %d values" % nm ** 3)
    ne = u.numel() // nm ** 3
    _check("prepared", prepared, _lib.lib().bkx_bk3_basis_size(p))
    _check("G", G, ne * 6 * nq ** 3, sdt)
    _check("u", u, ne * nm ** 3, sdt)
    if out is None:
        out = torch.empty(u.shape, dtype=torch.float64, device=u.device)
    _check("out", out, ne * nm ** 3)
    _apart(out, u)
    with _on(u):
        _lib.call(entry, p, ne, prepared.data_ptr(), G.data_ptr(), u.data_ptr(), out.data_ptr(), _stream(u))
    return out


_BK5_FORMS = {
    # form: (C entry point, storage dtype of u and G)
    "f64": ("bkx_bk5_apply_f64", torch.float64),
    "f32c": ("bkx_bk5_apply_f32c", torch.float32),
    "f32w": ("bkx_bk5_apply_f32w", torch.float32),
    "bf16s": ("bkx_bk5_apply_bf16s", torch.float32),
    "oz": ("bkx_bk5_apply_oz", torch.float32),
}


def bk5(p, dbasis, G, u, out=None, form="f64"):
    """BK5 collocated Laplacian v = D^T G D u per element.

    form "f64": u, G, v in FP64 (the drop-in);
         "f32c": u, G in FP32, float-float compensated arithmetic, v FP64;
         "f32w": u, G in FP32, FP64 arithmetic, v FP64;
         "bf16s": u, G in FP32, split-BF16 MFMA contractions accumulated in
                  FP32 (p = 7 only), v FP32;
         "oz": u, G in FP32, contractions from exact 8-bit digit products on
               BF16 MFMA with exact FP32 level sums (p = 5, 6, 7), v FP64."""
    entry, args, out = _bk5_call(p, dbasis, G, u, out, form)
    _launch(entry, args, u)
    return out


def _bk5_call(p, dbasis, G, u, out, form):
    """bk5's checks, as _bk1_call."""
    _order_ok(5, p)
    if form not in _BK5_FORMS:
        raise ValueError("unknown BK5 form %r (one of %s)" % (form, ", ".join(_BK5_FORMS)))
    entry, sdt = _BK5_FORMS[form]
    nq = p + 1
    if u.numel() % nq ** 3:
        raise ValueError("u must hold whole elements of %d values" % nq ** 3)
    ne = u.numel() // nq ** 3
    _check("dbasis", dbasis, nq * nq)
    _check("G", G, ne * 6 * nq ** 3, sdt)
    _check("u", u, ne * nq ** 3, sdt)
    odt = torch.float32 if form == "bf16s" else torch.float64
    if out is None:
        out = torch.empty(u.shape, dtype=odt, device=u.device)
    _check("out", out, ne * nq ** 3, odt)
    _apart(out, u)
    return entry, (p, ne, dbasis.data_ptr(), G.data_ptr(), u.data_ptr(), out.data_ptr()), out


def bind_bk1(p, basis, JxW, u, out=None, form="f64"):
    """bk1 with its arguments checked once: a Bound whose calls apply it."""
    entry, args, out = _bk1_call(p, basis, JxW, u, out, form)
    return Bound(entry, args, (basis, JxW, u, out), out)


def bind_bk3(p, basis, dbasis, G, u, out=None, form="f64"):
    """bk3 with its arguments checked once: a Bound whose calls apply it."""
    entry, args, out = _bk3_call(p, basis, dbasis, G, u, out, form)
    return Bound(entry, args, (basis, dbasis, G, u, out), out)


def bind_bk5(p, dbasis, G, u, out=None, form="f64"):
    """bk5 with its arguments checked once: a Bound whose calls apply it."""
    entry, args, out = _bk5_call(p, dbasis, G, u, out, form)
    return Bound(entry, args, (dbasis, G, u, out), out)


_KAPPA_FORMS = {"f64": torch.float64, "f32w": torch.float32, "f32c": torch.float32}


def _bk3_common(p, basis, dbasis, G, u, form="f64"):
    _order_ok(3, p)
    if form not in _KAPPA_FORMS:
        raise ValueError("unknown BK3-kappa form %r (one of %s)" % (form, ", ".join(_KAPPA_FORMS)))
    sdt = _KAPPA_FORMS[form]
    nm, nq = p + 1, p + 2
    if u.numel() % nm ** 3:
        raise ValueError("u must hold whole elements of %d values" % nm ** 3)
    ne = u.numel() // nm ** 3
    _check("basis", basis, nm * nq)
    _check("dbasis", dbasis, nq * nq)
    _check("G", G, ne * 6 * nq ** 3, sdt)
    _check("u", u, ne * nm ** 3, sdt)
    return ne, nm, nq, sdt


def bk3_kappa(p, basis, dbasis, G, kappa, u, out=None, form="f64"):
    """v = A(kappa) u: BK3 with each Gauss point's six factors scaled by
    kappa(e, p, q, r).  form "f64": every array FP64; "f32w": u, G and kappa
    stored in FP32, FP64 arithmetic; "f32c": the same FP32 storage with
    float-float compensated arithmetic; v FP64 in every form."""
    ne, nm, nq, sdt = _bk3_common(p, basis, dbasis, G, u, form)
    _check("kappa", kappa, ne * nq ** 3, sdt)
    if out is None:
        out = torch.empty(u.shape, dtype=torch.float64, device=u.device)
    _check("out", out, ne * nm ** 3)
    _apart(out, u)
    with _on(u):
        _lib.call("bkx_bk3_apply_kappa_" + form, p, ne, basis.data_ptr(), dbasis.data_ptr(), G.data_ptr(),
                  kappa.data_ptr(), u.data_ptr(), out.data_ptr(), _stream(u))
    return out


def bk3_adjoint(p, basis, dbasis, G, kappa, u, b, form="f64"):
    """For L = 1/2 |A(kappa) u - b|^2: returns (L, dL/du, dL/dkappa, loss per
    element), L as a 0-d device tensor summed in a fixed order.  form as for
    bk3_kappa (b in the storage type of u); the results are FP64."""
    ne, nm, nq, sdt = _bk3_common(p, basis, dbasis, G, u, form)
    _check("kappa", kappa, ne * nq ** 3, sdt)
    _check("b", b, ne * nm ** 3, sdt)
    dldu = torch.empty(u.shape, dtype=torch.float64, device=u.device)
    dldk = torch.empty(kappa.shape, dtype=torch.float64, device=u.device)
    loss_e = torch.empty(ne, dtype=torch.float64, device=u.device)
    L = torch.empty((), dtype=torch.float64, device=u.device)
    with _on(u):
        _lib.call("bkx_bk3_adjoint_" + form, p, ne, basis.data_ptr(), dbasis.data_ptr(), G.data_ptr(),
                  kappa.data_ptr(), u.data_ptr(), b.data_ptr(), dldu.data_ptr(), dldk.data_ptr(), loss_e.data_ptr(),
                  _stream(u))
        _lib.call("bkx_sum_f64", loss_e.data_ptr(), ne, L.data_ptr(), _stream(u))
    return L, dldu, dldk, loss_e


def bk3_adjoint_prepared(p, prepared, G, kappa, u, b, form="f64"):
    """bk3_adjoint with the bases in the even-odd form of bk3_basis; form
    "f64" or "f32w".  The results agree with bk3_adjoint's to rounding."""
    _order_ok(3, p)
    if form not in ("f64", "f32w"):
        raise ValueError("form must be f64 or f32w")
    sdt = torch.float64 if form == "f64" else torch.float32
    nm, nq = p + 1, p + 2
    if u.numel() % nm ** 3:
        raise ValueError("u must hold whole elements of %d values" % nm ** 3)
    ne = u.numel() // nm ** 3
    _check("prepared", prepared, _lib.lib().bkx_bk3_basis_size(p))
    _check("G", G, ne * 6 * nq ** 3, sdt)
    _check("kappa", kappa, ne * nq ** 3, sdt)
    _check("u", u, ne * nm ** 3, sdt)
    _check("b", b, ne * nm ** 3, sdt)
    dldu = torch.empty(u.shape, dtype=torch.float64, device=u.device)
    dldk = torch.empty(kappa.shape, dtype=torch.float64, device=u.device)
    loss_e = torch.empty(ne, dtype=torch.float64, device=u.device)
    L = torch.empty((), dtype=torch.float64, device=u.device)
    with _on(u):
        _lib.call("bkx_bk3_adjoint_prepared_" + form, p, ne, prepared.data_ptr(), G.data_ptr(), kappa.data_ptr(),
                  u.data_ptr(), b.data_ptr(), dldu.data_ptr(), dldk.data_ptr(), loss_e.data_ptr(), _stream(u))
        _lib.call("bkx_sum_f64", loss_e.data_ptr(), ne, L.data_ptr(), _stream(u))
    return L, dldu, dldk, loss_e


def bk3_vjp(p, basis, dbasis, G, kappa, u, w, form="f64"):
    """Vector-Jacobian product of v = A(kappa) u with cotangent w:
    (A(kappa) w, (grad w) . G (grad u)).  form as for bk3_kappa (w in the
    storage type of u); the results are FP64."""
    ne, nm, nq, sdt = _bk3_common(p, basis, dbasis, G, u, form)
    _check("kappa", kappa, ne * nq ** 3, sdt)
    _check("w", w, ne * nm ** 3, sdt)
    gu = torch.empty(u.shape, dtype=torch.float64, device=u.device)
    gk = torch.empty(kappa.shape, dtype=torch.float64, device=u.device)
    with _on(u):
        _lib.call("bkx_bk3_vjp_" + form, p, ne, basis.data_ptr(), dbasis.data_ptr(), G.data_ptr(), kappa.data_ptr(),
                  u.data_ptr(), w.data_ptr(), gu.data_ptr(), gk.data_ptr(), _stream(u))
    return gu, gk


def geo_bk3(p, basis, dbasis, qweights, coords, G=None, JxW=None):
    """Geometric factors at the q = p + 2 Gauss points from element nodal
    coordinates coords(e, 3, n, n, n) at the GLL nodes: G (E, 6, q, q, q)
    factor-major (BK3) and JxW (E, q, q, q) (BK1), both FP64; pass the output
    tensors, or None for a new one.  Returns (G, JxW)."""
    _order_ok(3, p)
    nm, nq = p + 1, p + 2
    if coords.numel() % (3 * nm ** 3):
        raise ValueError("coords must hold whole elements of 3 x %d values" % nm ** 3)
    ne = coords.numel() // (3 * nm ** 3)
    _check("basis", basis, nm * nq)
    _check("dbasis", dbasis, nq * nq)
    _check("qweights", qweights, nq)
    _check("coords", coords, ne * 3 * nm ** 3)
    if G is None:
        G = torch.empty((ne, 6, nq, nq, nq), dtype=torch.float64, device=coords.device)
    if JxW is None:
        JxW = torch.empty((ne, nq, nq, nq), dtype=torch.float64, device=coords.device)
    _check("G", G, ne * 6 * nq ** 3)
    _check("JxW", JxW, ne * nq ** 3)
    with _on(coords):
        _lib.call("bkx_geo_bk3_f64", p, ne, basis.data_ptr(), dbasis.data_ptr(), qweights.data_ptr(),
                  coords.data_ptr(), G.data_ptr(), JxW.data_ptr(), _stream(coords))
    return G, JxW


def geo_bk5(p, dbasis, qweights, coords, out=None):
    """BK5 geometric factors G (E, n, n, 6, n) interleaved at the n = p + 1
    GLL nodes from element nodal coordinates coords(e, 3, n, n, n), FP64."""
    _order_ok(5, p)
    n = p + 1
    if coords.numel() % (3 * n ** 3):
        raise ValueError("coords must hold whole elements of 3 x %d values" % n ** 3)
    ne = coords.numel() // (3 * n ** 3)
    _check("dbasis", dbasis, n * n)
    _check("qweights", qweights, n)
    _check("coords", coords, ne * 3 * n ** 3)
    if out is None:
        out = torch.empty((ne, n, n, 6, n), dtype=torch.float64, device=coords.device)
    _check("out", out, ne * 6 * n ** 3)
    with _on(coords):
        _lib.call("bkx_geo_bk5_f64", p, ne, dbasis.data_ptr(), qweights.data_ptr(), coords.data_ptr(),
                  out.data_ptr(), _stream(coords))
    return out


_LAYOUT_SUFFIX = {torch.float64: "f64", torch.float32: "f32"}


def bk5_make_G(p, G6i, F=6, layout=0, out=None):
    """BK5's interleaved six factors G6i (E, n, n, 6, n) in another layout
    (make_bk5_G, include/bk_harness.h:217-244): F = 6 or 9 factors per point,
    layout 0 factor-major (E, F, n, n, n), 1 interleaved (E, n, n, F, n), 2
    column (E, F, n, n, n) indexed (f, k, i, j).  FP64 or FP32, as G6i."""
    _order_ok(5, p)
    n = p + 1
    if G6i.dtype not in _LAYOUT_SUFFIX or G6i.numel() % (6 * n ** 3):
        raise ValueError("G6i must be FP64 or FP32 and hold whole elements of 6 x %d factors" % n ** 3)
    ne = G6i.numel() // (6 * n ** 3)
    _check("G6i", G6i, ne * 6 * n ** 3, G6i.dtype)
    if out is None:
        out = torch.empty(ne * F * n ** 3, dtype=G6i.dtype, device=G6i.device)
    _check("out", out, ne * F * n ** 3, G6i.dtype)
    with _on(G6i):
        _lib.call("bkx_bk5_make_G_" + _LAYOUT_SUFFIX[G6i.dtype], p, ne, G6i.data_ptr(), F, layout, out.data_ptr(),
                  _stream(G6i))
    return out


def bk5_interleave_G(p, G, F=6, layout=0, out=None):
    """The inverse of bk5_make_G: F = 6 or 9 factors in the given layout ->
    the interleaved six factors that bk5 reads."""
    _order_ok(5, p)
    n = p + 1
    if G.dtype not in _LAYOUT_SUFFIX or G.numel() % (F * n ** 3):
        raise ValueError("G must be FP64 or FP32 and hold whole elements of %d x %d factors" % (F, n ** 3))
    ne = G.numel() // (F * n ** 3)
    _check("G", G, ne * F * n ** 3, G.dtype)
    if out is None:
        out = torch.empty(ne * 6 * n ** 3, dtype=G.dtype, device=G.device)
    _check("out", out, ne * 6 * n ** 3, G.dtype)
    with _on(G):
        _lib.call("bkx_bk5_interleave_G_" + _LAYOUT_SUFFIX[G.dtype], p, ne, F, layout, G.data_ptr(), out.data_ptr(),
                  _stream(G))
    return out


def bk1_transpose_jxw(p, JxW, out=None):
    """BK1's JxW(e, p, q, r) -> JxW_t(e, r, q, p) (transpose_jxw,
    include/bk_harness.h:87-102).  FP64 or FP32, as JxW."""
    _order_ok(1, p)
    nq = p + 2
    if JxW.dtype not in _LAYOUT_SUFFIX or JxW.numel() % nq ** 3:
        raise ValueError("JxW must be FP64 or FP32 and hold whole elements of %d values" % nq ** 3)
    ne = JxW.numel() // nq ** 3
    _check("JxW", JxW, ne * nq ** 3, JxW.dtype)
    if out is None:
        out = torch.empty_like(JxW)
    _check("out", out, ne * nq ** 3, JxW.dtype)
    with _on(JxW):
        _lib.call("bkx_bk1_transpose_jxw_" + _LAYOUT_SUFFIX[JxW.dtype], p, ne, JxW.data_ptr(), out.data_ptr(),
                  _stream(JxW))
    return out
