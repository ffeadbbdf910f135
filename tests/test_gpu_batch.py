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


def test_batch_shared_plans_and_outputs(gpu):
    # entries that share a plan (its workspace) or an accumulated output run
    # in one graph branch, in batch order: a plan used twice gives both
    # results, and two plans adding into one output give the sum
    import torch
    ncols = 4
    kernel = pbb.exponential(0.5)
    ops = pbb.Operators(kernel, 1.0, 4, 4, 1e-7)
    pa = pbb.Plan(ops, uniform_points(6000, seed=80))
    pb = pbb.Plan(ops, uniform_points(6000, seed=81))
    pc = pbb.Plan(ops, clustered_points(5000, seed=82))
    g = torch.Generator(device="cuda").manual_seed(83)
    w = [torch.randn(ncols, n, dtype=torch.float64, device="cuda", generator=g) for n in (6000, 6000, 6000, 5000)]
    o_a1, o_a2 = (torch.zeros(ncols, 6000, dtype=torch.float64, device="cuda") for _ in range(2))
    o_sum = torch.zeros(ncols, 6000, dtype=torch.float64, device="cuda")
    o_c = torch.zeros(ncols, 5000, dtype=torch.float64, device="cuda")
    plans = [pa, pa, pa, pb, pc]
    ws = [w[0], w[1], w[2], w[1], w[3]]
    outs = [o_a1, o_a2, o_sum, o_sum, o_c]
    batch = pbb.Batch(plans, ws, outs, ncols, overwrite=False)
    e_a1, e_a2, e_a3 = pa.matvec(w[0]), pa.matvec(w[1]), pa.matvec(w[2])
    e_b, e_c = pb.matvec(w[1]), pc.matvec(w[3])
    torch.cuda.synchronize()
    for rep in range(3):
        for o in (o_a1, o_a2, o_sum, o_c):
            o.zero_()
        torch.cuda.synchronize()
        batch.launch()
        torch.cuda.synchronize()
        assert torch.equal(o_a1, e_a1) and torch.equal(o_a2, e_a2) and torch.equal(o_c, e_c)
        # the batch adds each result's parts into o_sum in turn: equal to the
        # sum to rounding (a race between branches would lose whole updates)
        err = float(torch.linalg.norm(o_sum - (e_a3 + e_b)) / torch.linalg.norm(e_a3 + e_b))
        assert err <= 1e-13, f"replay {rep}: the shared output is not the sum ({err:.2e})"


def test_batch_workspace_survives_eager_calls(gpu):
    # the batch owns its workspaces: eager calls on its plans with another
    # precision class and more columns (which regrow the plan's own
    # workspace) leave the captured graph valid
    import torch
    ncols = 4
    kernel = pbb.exponential(0.5)
    ops = pbb.Operators(kernel, 1.0, 4, 5, 1e-7)
    pa = pbb.Plan(ops, clustered_points(7000, seed=90))
    pb = pbb.Plan(ops, uniform_points(5000, seed=91))
    g = torch.Generator(device="cuda").manual_seed(92)
    wa = torch.randn(ncols, 7000, dtype=torch.float64, device="cuda", generator=g)
    wb = torch.randn(ncols, 5000, dtype=torch.float64, device="cuda", generator=g)
    oa, ob = torch.zeros_like(wa), torch.zeros_like(wb)
    batch = pbb.Batch([pa, pb], [wa, wb], [oa, ob], ncols, precision=pbb.capi.PREC_MIXED_F16X2)
    ea = pa.matvec(wa, precision=pbb.capi.PREC_MIXED_F16X2)
    eb = pb.matvec(wb, precision=pbb.capi.PREC_MIXED_F16X2)
    torch.cuda.synchronize()
    wide = torch.randn(3 * ncols, 7000, dtype=torch.float64, device="cuda", generator=g)
    for rep in range(2):
        pa.matvec(wide)                     # FP64, 12 columns: regrows pa's workspace
        pb.matvec_transpose(wb[:1].contiguous())
        torch.cuda.synchronize()
        batch.launch()
        torch.cuda.synchronize()
        assert torch.equal(oa, ea) and torch.equal(ob, eb), f"replay {rep}"


def test_batch_chained_entries(gpu):
    # entry 2 reads entry 0's output (A (A w)) through a second plan of the
    # same points: only the buffers link them, and they must run in order in
    # one branch, beside an unrelated entry
    import torch
    ncols = 2
    kernel = pbb.exponential(0.5)
    ops = pbb.Operators(kernel, 1.0, 4, 4, 1e-7)
    pa = pbb.Plan(ops, uniform_points(6000, seed=93))
    pa2 = pbb.Plan(ops, uniform_points(6000, seed=93))
    pc = pbb.Plan(ops, clustered_points(4000, seed=94))
    g = torch.Generator(device="cuda").manual_seed(95)
    w = torch.randn(ncols, 6000, dtype=torch.float64, device="cuda", generator=g)
    wc = torch.randn(ncols, 4000, dtype=torch.float64, device="cuda", generator=g)
    mid, fin, oc = torch.zeros_like(w), torch.zeros_like(w), torch.zeros_like(wc)
    batch = pbb.Batch([pa, pc, pa2], [w, wc, mid], [mid, oc, fin], ncols)
    e_mid = pa.matvec(w)
    e_fin = pa2.matvec(e_mid)
    e_c = pc.matvec(wc)
    torch.cuda.synchronize()
    for rep in range(3):
        batch.launch()
        torch.cuda.synchronize()
        assert torch.equal(mid, e_mid) and torch.equal(fin, e_fin) and torch.equal(oc, e_c), f"replay {rep}"


def test_wrappers_reject_bad_tensors(gpu):
    import torch
    ops = pbb.Operators(pbb.laplacian(), 1.0, 3, 4, 1e-5)
    p = pbb.Plan(ops, uniform_points(3000, seed=96))
    w = torch.randn(2, 3000, dtype=torch.float64, device="cuda")
    with pytest.raises(TypeError):
        p.matvec(w.float())
    with pytest.raises(ValueError):
        p.matvec(w, out=torch.zeros(3000, 2, dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError):
        p.matvec_transpose(w, out=torch.zeros(2, 3000, dtype=torch.float32, device="cuda"))
    with pytest.raises(TypeError):
        p.grad_positions(w, w.float())
    with pytest.raises(ValueError):
        p.grad_positions(w, w, grad=torch.zeros(3, 3000, dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError):
        p.grad_positions(w, w, grad=torch.zeros(3000, 3, dtype=torch.float32, device="cuda"))
    with pytest.raises(TypeError):
        pbb.Batch([p], [w.float()], [torch.zeros_like(w)], 2)
    with pytest.raises(TypeError):
        pbb.Batch([p], [w], [torch.zeros(2, 3000, dtype=torch.float64)], 2)


def test_256_clouds_forward_and_adjoint_in_one_graph(gpu):
    # 256 clouds, sizes log-uniform in [2e3, 2e5] (U[0,1)^3, seed 1000 + k),
    # tree level round(log8(N/60)) clamped to 3..6, orders 4..6 in turn,
    # exp(-r/0.5), 16 columns, eps 1e-7: the forward matvec AND the adjoint
    # of every cloud in one graph launch, FP64 and f16x2 (M2L tiles tuned
    # per operator set within 1e-6 of FP64); every replay equals the eager
    # calls bit for bit
    import torch
    ncl, nc = 256, 16
    kernel = pbb.exponential(0.5)
    sizes = np.exp(np.random.default_rng(999).uniform(np.log(2e3), np.log(2e5), ncl)).astype(int)
    ops, plans, sig, psi = {}, [], [], []
    for k in range(ncl):
        n = int(sizes[k])
        lvl = int(min(6, max(3, round(np.log(n / 60.0) / np.log(8.0)))))
        order = 4 + k % 3
        if (lvl, order) not in ops:
            ops[(lvl, order)] = pbb.Operators(kernel, 1.0, lvl, order, 1e-7)
        plans.append(pbb.Plan(ops[(lvl, order)], np.random.default_rng(1000 + k).random((n, 3))))
        g = torch.Generator(device="cuda").manual_seed(3000 + k)
        sig.append(torch.randn(nc, n, dtype=torch.float64, device="cuda", generator=g))
        psi.append(torch.randn(nc, n, dtype=torch.float64, device="cuda", generator=g))
    print(f"{ncl} clouds, {int(sizes.sum())} points, {len(ops)} operator sets", flush=True)
    for prec in (pbb.capi.PREC_FP64, pbb.capi.PREC_MIXED_F16X2):
        if prec != pbb.capi.PREC_FP64:
            tiles = pbb.Batch.tune(plans, sig, prec, max_rel_diff=1e-6)
            assert all(t["rel_diff"] <= 1e-6 for t in tiles.values()), tiles
        ent_plans = [p for p in plans for _ in (0, 1)]
        ent_w = [w for pair in zip(sig, psi) for w in pair]
        ent_out = [torch.zeros_like(w) for w in ent_w]
        batch = pbb.Batch(ent_plans, ent_w, ent_out, nc, precision=prec, transpose=[False, True] * ncl)
        eager = []
        for p, s, q in zip(plans, sig, psi):
            eager.append(p.matvec(s, precision=prec))
            eager.append(p.matvec_transpose(q, precision=prec))
        torch.cuda.synchronize()
        for rep in range(2):
            batch.launch()
            torch.cuda.synchronize()
            for i, (o, e) in enumerate(zip(ent_out, eager)):
                assert torch.equal(o, e), f"precision {prec} replay {rep} entry {i} differs from eager"
        print(f"precision {prec}: {2 * ncl} entries in one graph equal the eager calls", flush=True)
        del batch
