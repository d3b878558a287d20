"""Element restriction on torch tensors: the L-vector (one value per global
node) to the operators' E-vectors and back, through the C ABI
(bkx_restrict_*), and the assembled operators R^T A R built from them.

    R = Restriction(offsets, lsize)     # offsets: host int32 (E, n^3)
    uE = R.gather(uL);  vL = R.scatter(vE)
    vL = R.scatter(ops.bk5(p, D, G, R.gather(uL)))
    vL = R.scatter(R.bk5(p, D, G, uL))  # the same, gather fused into BK5

No reference counterpart: the reference applies its operators to E-vectors
only.  The scatter is deterministic (each node's contributions summed in
ascending E-vector order)."""
import ctypes

import numpy as np
import torch

from . import _lib
from .ops import _on, _stream


class Restriction:
    def __init__(self, offsets, lsize):
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int32))
        if off.ndim != 2:
            raise ValueError("offsets must be (nelmt, elemsize)")
        self.nelmt, self.elemsize = off.shape
        self.lsize = int(lsize)
        self._plan = ctypes.c_void_p()
        _lib.call("bkx_restrict_create", self.nelmt, self.elemsize, off.ctypes.data_as(ctypes.c_void_p), self.lsize,
                  ctypes.byref(self._plan))

    def close(self):
        if self._plan:
            _lib.call("bkx_restrict_destroy", self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, name, t, n):
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != torch.float64:
            raise TypeError("%s must be a float64 tensor on a HIP device" % name)
        if not t.is_contiguous() or t.numel() != n:
            raise ValueError("%s must be contiguous with %d values" % (name, n))

    def gather(self, uL, out=None):
        """uE = R uL (nelmt * elemsize values)."""
        self._check("uL", uL, self.lsize)
        if out is None:
            out = torch.empty(self.nelmt * self.elemsize, dtype=torch.float64, device=uL.device)
        self._check("out", out, self.nelmt * self.elemsize)
        with _on(uL):
            _lib.call("bkx_restrict_gather_f64", self._plan, uL.data_ptr(), out.data_ptr(),
                      _stream(uL))
        return out

    def scatter(self, vE, out=None):
        """vL = R^T vE (lsize values)."""
        self._check("vE", vE, self.nelmt * self.elemsize)
        if out is None:
            out = torch.empty(self.lsize, dtype=torch.float64, device=vE.device)
        self._check("out", out, self.lsize)
        with _on(vE):
            _lib.call("bkx_restrict_scatter_f64", self._plan, vE.data_ptr(), out.data_ptr(),
                      _stream(vE))
        return out

    def bk5(self, p, D, G, uL, out=None):
        """vE = A_E (R uL): BK5 (FP64) with the gather fused into the
        operator's loads of u, the same bits as ops.bk5(p, D, G,
        self.gather(uL)).  Elements of (p + 1)^3 GLL nodes."""
        n = p + 1
        if self.elemsize != n ** 3:
            raise ValueError("the restriction's elements have %d nodes, BK5 at p = %d needs %d"
                             % (self.elemsize, p, n ** 3))
        self._check("uL", uL, self.lsize)
        self._check("D", D, n * n)
        self._check("G", G, self.nelmt * 6 * n ** 3)
        if out is None:
            out = torch.empty(self.nelmt * self.elemsize, dtype=torch.float64, device=uL.device)
        self._check("out", out, self.nelmt * self.elemsize)
        with _on(uL):
            _lib.call("bkx_bk5_apply_restricted_f64", self._plan, p, D.data_ptr(), G.data_ptr(), uL.data_ptr(),
                      out.data_ptr(), _stream(uL))
        return out
