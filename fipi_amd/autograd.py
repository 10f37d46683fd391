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
        dev = mobility.device if torch.is_tensor(mobility) else theta.device
        # each gradient goes back on its input's device and dtype
        ctx.save_for_backward(
            torch.as_tensor(gm, dtype=torch.float64).to(**_like(mobility, theta)),
            torch.as_tensor(gl, dtype=torch.float64).to(**_like(lam, theta)),
            torch.from_numpy(np.asarray(gth, dtype=np.float64)).to(device=theta.device, dtype=theta.dtype))
        return torch.tensor(loss, dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, grad):
        gm, gl, gth = ctx.saved_tensors
        return ((grad * gm).to(gm.dtype), (grad * gl).to(gl.dtype), (grad.to(gth.device) * gth).to(gth.dtype),
                None)


def _like(x, fallback):
    """Device and dtype of x if it is a tensor, else of fallback."""
    t = x if torch.is_tensor(x) else fallback
    return {"device": t.device, "dtype": t.dtype}


def calibration_loss(mobility, lam, theta, problem: Adjoint):
    return CalibrationLoss.apply(mobility, lam, theta, problem)
