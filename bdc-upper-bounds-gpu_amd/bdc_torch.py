"""torch.autograd wrapper of the smoothed dual bound, for gradcheck only.

U_beta(theta) = beta^-1 ln sum_x exp(beta D(x || softmax theta)) and its
gradient come from the engine's hand-written adjoint (bdc_dual through the C
ABI); this Function only hands them to autograd so that
torch.autograd.gradcheck can compare the adjoint with finite differences.
theta is a float64 CPU tensor of length 2^k.
"""
import torch

import bdc_gpu


class DualBound(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, engine, beta, precision=bdc_gpu.PREC_FP64):
        u, g = engine.dual(theta.detach().cpu().numpy(), beta, precision)
        ctx.save_for_backward(torch.from_numpy(g).to(theta.dtype))
        return theta.new_tensor(u)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None


def dual_bound(theta, engine, beta, precision=bdc_gpu.PREC_FP64):
    return DualBound.apply(theta, engine, beta, precision)
