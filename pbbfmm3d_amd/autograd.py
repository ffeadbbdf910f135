"""torch.autograd.Function over the FMM matvec.

``fmm_matvec(xyz, weights, ops)`` returns A~(xyz) @ weights (weights as
(ncols, N) or (N,)); its backward gives
  d/dweights  through the transposed FMM (pbb_plan_matvec_transpose), and
  d/dxyz      through pbb_plan_grad_positions (derivatives of the Chebyshev
              interpolants and of the P2P kernel; box membership and the
              bbox centre are held fixed, as they are piecewise constant).
The tree is built from xyz at every forward call, unless a plan is given:
then its points are moved to xyz (pbb_plan_update_points, which fails if a
point leaves its leaf box) and its tree is reused.  Weights of any floating
dtype (FP32 included) are handed to the kernels as FP64, the output has the
weights' dtype and every gradient the dtype of its input; the arithmetic is
the precision mode's.
"""
from __future__ import annotations

import torch

from . import _capi as capi
from .core import Operators, Plan


class FMMMatvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xyz, weights, ops: Operators, precision: int, plan):
        pts = xyz.detach().to(torch.float64).cpu().numpy()
        if plan is None:
            plan = Plan(ops, pts)
        else:
            if plan.ops is not ops:
                raise ValueError("the plan was built with other operators")
            plan.update_points(pts)
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
        return gx, gw, None, None, None


def fmm_matvec(xyz: torch.Tensor, weights: torch.Tensor, ops: Operators,
               precision: int = capi.PREC_FP64, plan: Plan = None) -> torch.Tensor:
    """A~(xyz) @ weights with gradients for both arguments.  plan: a plan of
    ops whose tree is reused (its points are moved to xyz); it holds the
    points of its last forward, so run a backward before the next forward
    with the same plan."""
    return FMMMatvec.apply(xyz, weights, ops, precision, plan)
