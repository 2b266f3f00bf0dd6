"""torch.autograd view of the Chebyshev-coefficient adjoint.

``ChebForceMatchLoss`` evaluates the force-matching loss of one or more
device contexts (one per atomic configuration) at a vector of EAM Chebyshev
value coefficients and returns, in backward, the gradient the HIP adjoint
kernels computed (``comd_eam_cheb_grad``).  PyTorch only carries the 3n
parameters; every force, energy and gradient evaluation runs in
``libcomd_hip.so``.
"""
from __future__ import annotations

import numpy as np
import torch


class ChebForceMatchLoss(torch.autograd.Function):
    """loss(coef) = sum over configurations c of
    sum_i |f_i(coef) - t_i|^2 / (3 N_c) + w_e (E(coef) - E_t)^2."""

    @staticmethod
    def forward(ctx, coef, sims, targets, w_e):
        c = coef.detach().to("cpu", torch.float64).numpy().reshape(3, -1)
        total = 0.0
        grad = np.zeros_like(c)
        for sim, (f_t, e_t) in zip(sims, targets):
            loss, g = sim.cheb_grad(c, f_t, e_t, w_e)
            total += loss
            grad += g
        ctx.save_for_backward(torch.from_numpy(grad.reshape(coef.shape)).to(coef.device, coef.dtype))
        return coef.new_tensor(total)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None


def force_matching_loss(coef, sims, targets, w_e=1.0):
    """coef: tensor of 3n value coefficients (phi, rho, F); sims: Simulation
    objects holding the configurations; targets: (f_target, e_target) per
    simulation, f_target in the simulation's slot layout."""
    return ChebForceMatchLoss.apply(coef, list(sims), list(targets), float(w_e))
