"""torch.autograd.Function over the device adjoint (sw_adjoint_run).

`ShapeLoss.apply(log_compliance, rest_length, pressure, adjoint, target)`
returns the shape loss sum |x(end) - target|^2 as a float64 scalar tensor;
its backward returns the adjoint's gradients with respect to the three
parameter tensors (float64, definition order, on the CPU: they are small
host arrays handed to the C-ABI, while the solve runs on the GPU).
"""
from __future__ import annotations

import numpy as np
import torch

from . import Adjoint


class ShapeLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_compliance: torch.Tensor, rest_length: torch.Tensor, pressure: torch.Tensor,
                adjoint: Adjoint, target) -> torch.Tensor:
        def host(t: torch.Tensor) -> np.ndarray:
            return t.detach().to("cpu", torch.float64).contiguous().numpy()

        wanted = any(ctx.needs_input_grad[:3])
        out = adjoint.run(target, host(log_compliance), host(rest_length), host(pressure), gradients=wanted)
        if wanted:
            grads = [torch.from_numpy(out[k]) for k in ("grad_log_compliance", "grad_rest_length", "grad_pressure")]
            ctx.save_for_backward(*grads)
        return torch.tensor(out["loss"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_loss: torch.Tensor):
        g_lc, g_rest, g_p = ctx.saved_tensors
        return grad_loss * g_lc, grad_loss * g_rest, grad_loss * g_p, None, None


def shape_loss(log_compliance: torch.Tensor, rest_length: torch.Tensor, pressure: torch.Tensor, adjoint: Adjoint,
               target) -> torch.Tensor:
    return ShapeLoss.apply(log_compliance, rest_length, pressure, adjoint, target)
