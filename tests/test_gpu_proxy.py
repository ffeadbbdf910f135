"""Proxy passes of the mixed modes: at the levels whose re-compressed ranks
are at most 32 the proxies are formed straight from the points (P = VT' S
summed point by point) and the M2L results are evaluated straight at the
points (value = (U'^T S_n(x))^T Pf), without the levels' n^3-coefficient
expansions.  Checked against the expansion route on the same plan and
operators (PBB_PLAN_EXPANSION_ROUTE: the same sums in another order, so FP32
rounding apart) and against the FP64 path, on point sets whose boxes
hold from one point to thousands (passes carried across, split upward work
items), with 1 to 40 columns, separate targets and sources, and the adjoint.
"""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc
from tests._helpers import clustered_points, weights

pytestmark = pytest.mark.gpu

KERNEL = pbb.gaussian(0.3)
LEVEL, ORDER, EPS = 6, 4, 1e-6  # re-compressed ranks of levels 5 and 6: 32 x 32 and 16 x 16


def _ops():
    ops = pbb.Operators(KERNEL, 1.0, LEVEL, ORDER, EPS)
    ranks = ops.mixed_ranks()
    # the proxy passes run at the levels whose ranks fit 32 (and every deeper
    # one); level 6 keeps 16 directions each side (ks stored padded to one
    # 32-deep chunk), so its f16 x2 M2L takes 16-deep steps
    assert [r[:2] for r in ranks[-2:]] == [(32, 32), (16, 16)], ranks
    assert all(r[0] > 32 for r in ranks[:-2])
    return ops


def _mv(plan, w, ncols, prec, transpose=False):
    import torch
    n_in = plan.ntargets if transpose else plan.nsources
    wd = torch.tensor(w.reshape(ncols, n_in), dtype=torch.float64, device="cuda:0")
    out = plan.matvec_transpose(wd, precision=prec) if transpose else plan.matvec(wd, precision=prec)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(-1)


@pytest.mark.parametrize("ncols", [1, 7, 32, 40])
def test_proxy_passes_equal_expansion_route(gpu, ncols):
    # tight clusters: level-5/6 boxes of several thousand points next to
    # boxes of one point
    N = 120000
    xyz = clustered_points(N, seed=71, sigma=0.015)
    w = weights(N, ncols, seed=72, dist="normal")
    ops = _ops()
    plan = pbb.Plan(ops, xyz)
    ref = pbb.Plan(ops, xyz, expansion_route=True)
    info = plan.info()
    print(f"case ncols={ncols}: boxes per level {info['boxes'][:LEVEL + 1]}, "
          f"largest leaf {info['max_leaf_points']}", flush=True)
    assert info["max_leaf_points"] > 2 * 128  # boxes span several passes
    f64 = _mv(plan, w, ncols, pbb.capi.PREC_FP64)
    assert np.array_equal(f64, _mv(ref, w, ncols, pbb.capi.PREC_FP64))  # FP64 has no proxy passes
    for prec in (pbb.capi.PREC_MIXED_F16X2, pbb.capi.PREC_MIXED):
        px = _mv(plan, w, ncols, prec)
        ex = _mv(ref, w, ncols, prec)
        d_route, d_px, d_ex = orc.rel_err(px, ex), orc.rel_err(px, f64), orc.rel_err(ex, f64)
        print(f"case ncols={ncols} prec={prec}: proxy vs expansion route {d_route:.2e}, "
              f"vs FP64 {d_px:.2e} (expansion route {d_ex:.2e})", flush=True)
        # signed weights: the boxes' sums cancel, which a one-column relative
        # norm shows most.  The two routes round in FP32 at different places
        # (measured 3.8e-7 apart at one column); the proxy route is as close
        # to FP64 as the expansion route
        assert d_route <= 5e-7
        assert d_px <= 1.25 * d_ex + 1e-7
        if ncols >= 32:
            assert d_px <= 1e-6


def test_proxy_passes_separate_sets_and_adjoint(gpu):
    rng = np.random.default_rng(73)
    tgt = clustered_points(50000, seed=74, sigma=0.02)  # in the unit cube
    src = 0.1 + 0.8 * rng.random((30000, 3))  # inside the targets' root cube
    nc = 16
    ops = _ops()
    plan = pbb.Plan(ops, tgt, sources=src)
    ref = pbb.Plan(ops, tgt, sources=src, expansion_route=True)
    assert plan.info()["outside_points"] == 0
    w = weights(src.shape[0], nc, seed=75, dist="normal")
    g = weights(tgt.shape[0], nc, seed=76, dist="normal")
    for transpose, wt in ((False, w), (True, g)):
        f64 = _mv(plan, wt, nc, pbb.capi.PREC_FP64, transpose)
        px = _mv(plan, wt, nc, pbb.capi.PREC_MIXED_F16X2, transpose)
        ex = _mv(ref, wt, nc, pbb.capi.PREC_MIXED_F16X2, transpose)
        d_route, d_px = orc.rel_err(px, ex), orc.rel_err(px, f64)
        print(f"case separate sets transpose={transpose}: proxy vs expansion route {d_route:.2e}, "
              f"vs FP64 {d_px:.2e}", flush=True)
        assert d_route <= 5e-7 and d_px <= 1e-6


def test_outside_points_keep_the_expansion_route(gpu):
    # points outside the root cube are extrapolated (as the reference does);
    # their plans keep the expansion route, so the mixed result stays as close
    # to FP64 as with PBB_PLAN_EXPANSION_ROUTE, bit for bit the same sums
    rng = np.random.default_rng(77)
    tgt = clustered_points(20000, seed=78, sigma=0.02)
    src = rng.random((10000, 3)) - 0.3  # partly outside
    ops = _ops()
    plan = pbb.Plan(ops, tgt, sources=src)
    ref = pbb.Plan(ops, tgt, sources=src, expansion_route=True)
    assert plan.info()["outside_points"] > 0
    w = weights(src.shape[0], 8, seed=79)
    a = _mv(plan, w, 8, pbb.capi.PREC_MIXED_F16X2)
    b = _mv(ref, w, 8, pbb.capi.PREC_MIXED_F16X2)
    assert np.array_equal(a, b)
