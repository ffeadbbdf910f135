"""torch.autograd.Function over the FMM matvec.

``fmm_matvec(xyz, weights, ops)`` returns A~(xyz) @ weights (weights as
(ncols, N) or (N,)); its backward gives
  d/dweights  through the transposed FMM (pbb_plan_matvec_transpose), and
  d/dxyz      through pbb_plan_grad_positions (derivatives of the Chebyshev
              interpolants and of the P2P kernel; box membership and the
              bbox centre are held fixed, as they are piecewise constant).
The tree is built from xyz at every forward call.  Inputs of any floating
dtype are computed in FP64 and gradients are returned in the input dtype.
"""
from __future__ import annotations

import torch

from . import _capi as capi
from .core import Operators, Plan


class FMMMatvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, weights, ops: Operators, precision: int):
        plan = Plan(ops, xyz.detach().to(torch.float64).cpu().numpy())
        w64 = weights.detach().to(device="cuda", dtype=torch.float64).contiguous()
        out = plan.matvec(w64, precision=precision)
        ctx.plan = plan
        ctx.precision = precision
        ctx.xyz_meta = (xyz.dtype, xyz.device)
        ctx.save_for_backward(w64)
        return out.to(weights.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (w64,) = ctx.saved_tensors
        g64 = grad_out.detach().to(device="cuda", dtype=torch.float64).contiguous()
        gx = gw = None
        if ctx.needs_input_grad[1]:
            gw = ctx.plan.matvec_transpose(g64, precision=ctx.precision).to(grad_out.dtype)
        if ctx.needs_input_grad[0]:
            dt, dev = ctx.xyz_meta
            gx = ctx.plan.grad_positions(w64, g64, precision=ctx.precision).to(device=dev, dtype=dt)
        return gx, gw, None, None


def fmm_matvec(xyz: torch.Tensor, weights: torch.Tensor, ops: Operators,
               precision: int = capi.PREC_FP64) -> torch.Tensor:
    """A~(xyz) @ weights with gradients for both arguments."""
    return FMMMatvec.apply(xyz, weights, ops, precision)
