"""Batches of uneven clouds in one HIP graph: every replay (with eager work
and synchronisations between replays) must equal each plan's own eager
matvec, forward and transposed, FP64 and mixed."""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc
from tests._helpers import clustered_points, oracle_ops_from, uniform_points

pytestmark = pytest.mark.gpu


def _clouds(k, seed):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(2000), np.log(40000), k)).astype(int)
    clouds = []
    for i, n in enumerate(sizes):
        xyz = clustered_points(n, seed=1000 + i) if i % 2 else uniform_points(n, seed=1000 + i)
        level = int(np.clip(np.round(np.log(n / 60) / np.log(8)), 3, 5))
        order = 4 + i % 3
        clouds.append((xyz, level, order))
    return clouds


@pytest.mark.parametrize("prec,transpose", [(pbb.capi.PREC_FP64, False), (pbb.capi.PREC_FP64, True),
                                            (pbb.capi.PREC_MIXED, False)],
                         ids=["fp64", "fp64-transpose", "mixed"])
def test_batch_replays_match_eager(gpu, prec, transpose):
    import torch
    ncols = 16
    kernel = pbb.exponential(0.5)
    ops_by_order = {n: pbb.Operators(kernel, 1.0, 5, n, 1e-7) for n in (4, 5, 6)}
    clouds = _clouds(10, seed=60)
    plans, ws, outs = [], [], []
    for i, (xyz, level, order) in enumerate(clouds):
        ops = pbb.Operators(kernel, 1.0, level, order, 1e-7) if level != 5 else ops_by_order[order]
        p = pbb.Plan(ops, xyz)
        plans.append(p)
        g = torch.Generator(device="cuda").manual_seed(70 + i)
        ws.append(torch.randn(ncols, len(xyz), dtype=torch.float64, device="cuda", generator=g))
        outs.append(torch.zeros(ncols, len(xyz), dtype=torch.float64, device="cuda"))
    batch = pbb.Batch(plans, ws, outs, ncols, precision=prec, transpose=transpose)
    eager = [(p.matvec_transpose(w, precision=prec) if transpose else p.matvec(w, precision=prec))
             for p, w in zip(plans, ws)]
    torch.cuda.synchronize()
    for rep in range(3):
        batch.launch()
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert torch.equal(o, e), f"replay {rep} differs from the eager matvec"
        # eager work between replays, on the same buffers
        for w in ws:
            w.mul_(1.0)
        torch.cuda.synchronize()
    if not transpose and prec == pbb.capi.PREC_FP64:
        xyz, level, order = clouds[0]
        ref = orc.matvec(oracle_ops_from(plans[0].ops, kernel), xyz, ws[0].cpu().numpy().reshape(-1), ncols)
        assert orc.rel_err(outs[0].cpu().numpy().reshape(-1), ref) <= 1e-10
