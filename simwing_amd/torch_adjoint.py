"""torch.autograd.Function over the device adjoint (sw_adjoint_run).

`ShapeLoss.apply(log_compliance, rest_length, pressure, adjoint, target)`
returns the shape loss sum |x(end) - target|^2 as a float64 scalar tensor
on the device of `log_compliance`; its backward returns the adjoint's
gradients with respect to the three parameter tensors (float64, definition
order), each on the device its parameter tensor lives on. The parameters
travel to the C-ABI as small host arrays; the solve runs on the GPU.
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
            params = (log_compliance, rest_length, pressure)
            keys = ("grad_log_compliance", "grad_rest_length", "grad_pressure")
            ctx.save_for_backward(*[torch.from_numpy(out[k]).to(p.device) for k, p in zip(keys, params)])
        return torch.tensor(out["loss"], dtype=torch.float64, device=log_compliance.device)

    @staticmethod
    def backward(ctx, grad_loss: torch.Tensor):
        g_lc, g_rest, g_p = ctx.saved_tensors
        return (grad_loss.to(g_lc.device) * g_lc, grad_loss.to(g_rest.device) * g_rest,
                grad_loss.to(g_p.device) * g_p, None, None)


def shape_loss(log_compliance: torch.Tensor, rest_length: torch.Tensor, pressure: torch.Tensor, adjoint: Adjoint,
               target) -> torch.Tensor:
    return ShapeLoss.apply(log_compliance, rest_length, pressure, adjoint, target)
