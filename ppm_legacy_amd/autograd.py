"""torch.autograd wrapper of P2M on a fully periodic mesh (compact layout).

    field = p2m(xp, up, mesh, kernel=KERNEL_MP4, prec=PREC_F64)

field has shape (N_z, N_y, N_x, lda) (unique nodes).  The backward pass runs
the hand-written adjoint kernels (ppmg_p2m_adjoint): the strength gradient is
the M2P transpose of P2M, the position gradient uses the M'4 (or BSP2)
derivative stencil; both reuse the forward's sort permutation and cell keys
kept in the plan.
"""
import torch

from . import core
from ._lib import KERNEL_MP4, PREC_F64


class P2MFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xp, up, mesh, kernel, prec):
        lda = up.shape[1]
        field = torch.empty(mesh.compact_shape(lda), dtype=core.VAL_DTYPE[prec], device=up.device)
        plan = core.p2m(mesh, xp.contiguous(), up.contiguous(), field, [0] * mesh.dim, kernel, prec)
        ctx.save_for_backward(xp, up)
        ctx.mesh, ctx.kernel, ctx.prec, ctx.plan = mesh, kernel, prec, plan
        return field

    @staticmethod
    def backward(ctx, gfield):
        xp, up = ctx.saved_tensors
        want_x, want_u = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (want_x or want_u):
            return None, None, None, None, None
        g = gfield.to(core.VAL_DTYPE[ctx.prec]).contiguous()
        gu, gx = core.p2m_adjoint(ctx.mesh, xp.contiguous(), up.contiguous(), g, [0] * ctx.mesh.dim, ctx.plan,
                                  ctx.kernel, ctx.prec, want_u=want_u, want_x=want_x)
        if gx is not None:
            gx = gx.to(xp.dtype)
        return gx, gu, None, None, None


def p2m(xp, up, mesh, kernel=KERNEL_MP4, prec=PREC_F64):
    if not mesh.periodic:
        raise ValueError("the autograd P2M needs a fully periodic mesh")
    return P2MFunction.apply(xp, up, mesh, kernel, prec)
