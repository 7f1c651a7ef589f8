"""The hand-written adjoint of the discrete loss behind torch autograd.

`discrete_loss(op, *fields)` returns loss = sum_e mean(f_e^2) of an
`odil_amd.operators` operator as a differentiable torch scalar. Forward
runs the fused loss + adjoint kernel once (`odil_<op>_loss_grad`) and keeps
its gradients; backward scales them by the incoming gradient. This replaces
the TensorFlow tape gradient of `Problem._eval_loss` (util_op.py:628-654)
for torch optimisers and any torch code composed around the loss. Frozen
(stop-gradient) paths of the operator are excluded exactly as in the
reference's tape.
"""

import torch


class DiscreteLoss(torch.autograd.Function):
    """loss = sum_e mean(f_e(fields)^2) with the kernel adjoint."""

    @staticmethod
    def forward(ctx, op, *fields):
        loss, grads, _ = op.loss_grad([f.detach().contiguous()
                                       for f in fields])
        ctx.save_for_backward(*grads)
        return loss.to(fields[0].dtype)

    @staticmethod
    def backward(ctx, gout):
        grads = ctx.saved_tensors
        return (None, ) + tuple(g * gout.to(g.dtype) for g in grads)


def discrete_loss(op, *fields):
    """Differentiable discrete loss of `op` at `fields` (device tensors)."""
    return DiscreteLoss.apply(op, *fields)


class InverseWaveLoss(torch.autograd.Function):
    """Loss of `operators.WaveInverse` in (u, c) with its kernel adjoint."""

    @staticmethod
    def forward(ctx, op, u, c):
        loss, gu, gc, _ = op.loss_grad(u.detach().contiguous(),
                                       c.detach().contiguous())
        ctx.save_for_backward(gu, gc)
        return loss.to(u.dtype)

    @staticmethod
    def backward(ctx, gout):
        gu, gc = ctx.saved_tensors
        g = gout.to(gu.dtype)
        return None, gu * g, gc * g
