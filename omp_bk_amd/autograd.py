"""torch.autograd.Function around the BK3 operator with a coefficient.

    v = Bk3Kappa.apply(u, kappa, p, basis, dbasis, G)

runs A(kappa) u on the device; its backward is the hand-written adjoint
kernel (bkx_bk3_vjp_f64): grad_u = A(kappa) grad_v (A is symmetric) and
grad_kappa = (grad grad_v) . G (grad u) at every Gauss point.  The basis
matrices and G are treated as constants (no gradient).  A thin test wrapper:
the kernels and their checks live in the C ABI."""
import torch

from . import ops


class Bk3Kappa(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, kappa, p, basis, dbasis, G):
        u = u.contiguous()
        kappa = kappa.contiguous()
        ctx.p = p
        ctx.save_for_backward(u, kappa, basis, dbasis, G)
        return ops.bk3_kappa(p, basis, dbasis, G, kappa, u)

    @staticmethod
    def backward(ctx, grad_v):
        u, kappa, basis, dbasis, G = ctx.saved_tensors
        gu, gk = ops.bk3_vjp(ctx.p, basis, dbasis, G, kappa, u, grad_v.contiguous())
        return gu, gk, None, None, None, None


def bk3_kappa(u, kappa, p, basis, dbasis, G):
    return Bk3Kappa.apply(u, kappa, p, basis, dbasis, G)
