"""torch.autograd.Function wrapper of the calibration adjoint (a thin test
harness: the forward run and the adjoint sweep are libfipi's; torch only
carries the three parameter tensors and their gradients)."""
from __future__ import annotations

import numpy as np
import torch

from . import Adjoint


class CalibrationLoss(torch.autograd.Function):
    """loss = sum (phi_T(mobility, lambda, theta) - phi_target)^2 for a
    fipi_amd.Adjoint problem; backward returns the adjoint gradients."""

    @staticmethod
    def forward(ctx, mobility, lam, theta, problem: Adjoint):
        th = theta.detach().cpu().double().numpy()
        loss, gm, gl, gth = problem.gradient(float(mobility), float(lam), th)
        ctx.save_for_backward(torch.tensor(gm, dtype=torch.float64), torch.tensor(gl, dtype=torch.float64),
                              torch.from_numpy(np.asarray(gth, dtype=np.float64)))
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        gm, gl, gth = ctx.saved_tensors
        return grad * gm, grad * gl, grad * gth, None


def calibration_loss(mobility, lam, theta, problem: Adjoint):
    return CalibrationLoss.apply(mobility, lam, theta, problem)
