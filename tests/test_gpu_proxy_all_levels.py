"""Mixed modes whose re-compressed ranks fit 32 at every level (a smooth
kernel at a low order): the proxy passes then run at every level from 2
down, with no expansion level above them (no P2M / M2M / L2L / L2P on S and
F at all).  Checked against PBB_PLAN_EXPANSION_ROUTE (S and F at every
level, the same sums in another order) and against the FP64 path, forward
and adjoint.
"""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc
from tests._helpers import clustered_points, weights

pytestmark = pytest.mark.gpu

KERNEL = pbb.gaussian(1.0)
LEVEL, ORDER, EPS = 4, 4, 1e-6  # kept ranks 32 x 32 at levels 2..4, reference ranks 32 / 32


def _mv(plan, w, ncols, prec, transpose=False):
    import torch
    n_in = plan.ntargets if transpose else plan.nsources
    wd = torch.tensor(w.reshape(ncols, n_in), dtype=torch.float64, device="cuda:0")
    out = plan.matvec_transpose(wd, precision=prec) if transpose else plan.matvec(wd, precision=prec)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(-1)


def test_proxy_passes_at_every_level(gpu):
    ops = pbb.Operators(KERNEL, 1.0, LEVEL, ORDER, EPS)
    ranks = ops.mixed_ranks()
    # every level's kept ranks fit the proxy passes' 32 rows, and the
    # reference rank r_f holds a 32-float Pf row: levels 2..4 all run them
    assert len(ranks) == LEVEL - 1 and all(r[0] <= 32 and r[1] <= 32 for r in ranks), ranks
    assert ops.rank_f >= 32, ops.rank_f
    N, nc = 60000, 8
    xyz = clustered_points(N, seed=81, sigma=0.02)
    plan = pbb.Plan(ops, xyz)
    ref = pbb.Plan(ops, xyz, expansion_route=True)
    print(f"case proxy passes at every level: ranks {ranks}, boxes {plan.info()['boxes'][:LEVEL + 1]}",
          flush=True)
    w = weights(N, nc, seed=82, dist="normal")
    for transpose in (False, True):
        f64 = _mv(plan, w, nc, pbb.capi.PREC_FP64, transpose)
        for prec in (pbb.capi.PREC_MIXED_F16X2, pbb.capi.PREC_MIXED):
            px = _mv(plan, w, nc, prec, transpose)
            ex = _mv(ref, w, nc, prec, transpose)
            d_route, d_px, d_ex = orc.rel_err(px, ex), orc.rel_err(px, f64), orc.rel_err(ex, f64)
            print(f"case transpose={transpose} prec={prec}: proxy vs expansion route {d_route:.2e}, "
                  f"vs FP64 {d_px:.2e} (expansion route {d_ex:.2e})", flush=True)
            assert d_route <= 5e-7
            assert d_px <= 1.25 * d_ex + 1e-7
            assert d_px <= 1e-6
