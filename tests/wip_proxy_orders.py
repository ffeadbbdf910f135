"""Proxy passes at orders 5, 7 and 8 (the kernels are templates over the
order; orders 4 and 6 are covered elsewhere): at level 5 of a Gaussian at
eps 1e-7 the re-compressed ranks of the deepest level fit 32, so that level
runs the proxy passes.  Checked against PBB_PLAN_EXPANSION_ROUTE and the
FP64 path, forward and adjoint.
"""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc
from tests._helpers import clustered_points, weights

pytestmark = pytest.mark.gpu

KERNEL = pbb.gaussian(0.3)
LEVEL, EPS = 5, 1e-7  # level 5 keeps 32 x 32 at orders 5, 7 and 8


def _mv(plan, w, ncols, prec, transpose=False):
    import torch
    n_in = plan.ntargets if transpose else plan.nsources
    wd = torch.tensor(w.reshape(ncols, n_in), dtype=torch.float64, device="cuda:0")
    out = plan.matvec_transpose(wd, precision=prec) if transpose else plan.matvec(wd, precision=prec)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(-1)


@pytest.mark.parametrize("order", [5, 7, 8])
def test_proxy_passes_by_order(gpu, order):
    ops = pbb.Operators(KERNEL, 1.0, LEVEL, order, EPS)
    ranks = ops.mixed_ranks()
    assert ranks[-1][0] <= 32 and ranks[-1][1] <= 32 and ops.rank_f >= 32, ranks
    N, nc = 40000, 8
    xyz = clustered_points(N, seed=91 + order, sigma=0.02)
    plan = pbb.Plan(ops, xyz)
    ref = pbb.Plan(ops, xyz, expansion_route=True)
    print(f"case order={order}: ranks {[r[:2] for r in ranks]}", flush=True)
    w = weights(N, nc, seed=92 + order, dist="normal")
    for transpose in (False, True):
        f64 = _mv(plan, w, nc, pbb.capi.PREC_FP64, transpose)
        px = _mv(plan, w, nc, pbb.capi.PREC_MIXED_F16X2, transpose)
        ex = _mv(ref, w, nc, pbb.capi.PREC_MIXED_F16X2, transpose)
        d_route, d_px, d_ex = orc.rel_err(px, ex), orc.rel_err(px, f64), orc.rel_err(ex, f64)
        print(f"case order={order} transpose={transpose}: proxy vs expansion route {d_route:.2e}, "
              f"vs FP64 {d_px:.2e} (expansion route {d_ex:.2e})", flush=True)
        assert d_route <= 5e-7
        assert d_px <= 1.25 * d_ex + 1e-7
