"""Separate targets and sources (pbb_plan_create_ts, pbb_matvec_ts) and the
plan options around them, on the GPU through the C ABI.

The oracle's matvec_ts (one tree over [targets; sources] centred on the
targets' bbox, as compute.hpp:92-127 centres on the targets, weights on the
sources) is itself pinned against the direct sum (tests/test_oracle.py);
here the FP64 GPU path must equal it to 1e-10, the transposed FMM must be
its exact adjoint, and the mixed modes must stay within 1e-6 of FP64.
"""
import threading

import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc
from tests._helpers import clustered_points, gpu_direct, oracle_ops_from, uniform_points, weights

PARITY = 1e-10

pytestmark = pytest.mark.gpu


def _sets(nt, ns, seed):
    # targets clustered; sources uniform over a slightly different box (a
    # few of them outside the targets' root cube)
    return clustered_points(nt, seed=seed), 0.95 * uniform_points(ns, seed=seed + 1) + 0.04


@pytest.mark.parametrize("kernel,level,n,eps,host_tree", [
    (pbb.laplacian(), 4, 4, 1e-5, False),
    (pbb.gaussian(0.3), 4, 5, 1e-7, False),
    (pbb.laplacian(), 5, 6, 1e-8, True),
], ids=["laplacian-l4", "gaussian-l4", "laplacian-l5-host-tree"])
def test_separate_sets_vs_oracle(gpu, kernel, level, n, eps, host_tree):
    nt, ns, ncols = 7000, 11000, 3
    print(f"case kernel={kernel.id} level={level} n={n} nt={nt} ns={ns}", flush=True)
    tgt, src = _sets(nt, ns, 80)
    w = weights(ns, ncols, seed=82)
    ops = pbb.Operators(kernel, 1.0, level, n, eps)
    plan = pbb.Plan(ops, tgt, sources=src, host_tree=host_tree)
    info = plan.info()
    assert (info["ntargets"], info["nsources"], info["npoints"]) == (nt, ns, nt + ns)
    import torch
    wd = torch.tensor(w.reshape(ncols, ns), device="cuda")
    got = plan.matvec(wd)
    assert got.shape == (ncols, nt)
    got = got.cpu().numpy().reshape(-1)
    ref = orc.matvec_ts(oracle_ops_from(ops, kernel), tgt, src, w, ncols)
    assert orc.rel_err(got, ref) <= PARITY
    # and against the exact sums (approximation level)
    sample = np.arange(0, nt, 37)
    d = gpu_direct(kernel, tgt[sample], src, w, ncols)
    rows = (sample[None, :] + nt * np.arange(ncols)[:, None]).ravel()
    assert orc.rel_err(got[rows], d) < 1e-3


def test_gpu_and_host_trees_agree_on_separate_sets(gpu):
    tgt, src = _sets(5000, 9000, 84)
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 5, 4, 1e-6)
    g = pbb.Plan(ops, tgt, sources=src)
    h = pbb.Plan(ops, tgt, sources=src, host_tree=True)
    assert g.info() == h.info()
    import torch
    w = torch.tensor(weights(9000, 2, seed=85).reshape(2, 9000), device="cuda")
    assert torch.equal(g.matvec(w), h.matvec(w))


def test_transpose_is_the_adjoint_and_mixed_modes(gpu):
    # <psi, A sigma> = <A^T psi, sigma> to rounding (the transposed FMM
    # applies the transpose of the same approximant); the transposed output
    # lives on the sources and matches the direct sum at approximation level
    import torch
    nt, ns, ncols = 6000, 9000, 2
    tgt, src = _sets(nt, ns, 86)
    kernel = pbb.gaussian(0.4)
    ops = pbb.Operators(kernel, 1.0, 4, 6, 1e-8)
    plan = pbb.Plan(ops, tgt, sources=src)
    rng = np.random.default_rng(87)
    sigma = torch.tensor(rng.standard_normal((ncols, ns)), device="cuda")
    psi = torch.tensor(rng.standard_normal((ncols, nt)), device="cuda")
    a_sigma = plan.matvec(sigma)
    at_psi = plan.matvec_transpose(psi)
    assert at_psi.shape == (ncols, ns)
    lhs = float((psi * a_sigma).sum())
    rhs = float((at_psi * sigma).sum())
    assert abs(lhs - rhs) <= 1e-11 * float(torch.linalg.norm(psi) * torch.linalg.norm(a_sigma))
    sample = np.arange(0, ns, 53)
    d = gpu_direct(kernel, src[sample], tgt, psi, ncols)
    rows = (sample[None, :] + ns * np.arange(ncols)[:, None]).ravel()
    assert orc.rel_err(at_psi.cpu().numpy().reshape(-1)[rows], d) < 1e-5
    for prec in (pbb.capi.PREC_MIXED_F16X2, pbb.capi.PREC_MIXED):
        # the mixed modes' 1e-6 bar is the tuner's bound (tuned on the forward FMM)
        plan.tune_m2l(sigma, prec, 1e-6)
        m = plan.matvec(sigma, precision=prec)
        assert float(torch.linalg.norm(m - a_sigma) / torch.linalg.norm(a_sigma)) <= 1e-6
        mt = plan.matvec_transpose(psi, precision=prec)
        assert float(torch.linalg.norm(mt - at_psi) / torch.linalg.norm(at_psi)) <= 2e-6


def test_same_points_as_separate_sets(gpu):
    # sources equal to the targets, passed as separate sets: every point is
    # in the tree twice (once without weight) and the sums are the same terms
    import torch
    xyz = clustered_points(8000, seed=88)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 4, 5, 1e-7)
    one = pbb.Plan(ops, xyz)
    two = pbb.Plan(ops, xyz, sources=xyz)
    w = torch.tensor(weights(8000, 2, seed=89).reshape(2, 8000), device="cuda")
    a, b = one.matvec(w), two.matvec(w)
    assert float(torch.linalg.norm(a - b) / torch.linalg.norm(a)) <= 1e-12


def test_host_one_shot_and_batch_with_separate_sets(gpu):
    import torch
    nt, ns, ncols = 3000, 5000, 2
    tgt, src = _sets(nt, ns, 90)
    w = weights(ns, ncols, seed=91)
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 4, 4, 1e-6)
    ref = orc.matvec_ts(oracle_ops_from(ops, pbb.gaussian(0.3)), tgt, src, w, ncols)
    out = np.full(nt * ncols, 0.25)
    pbb.matvec_host(ops, tgt, w, ncols, out, sources=src)  # pbb_matvec_ts, accumulates
    assert orc.rel_err(out - 0.25, ref) <= PARITY
    # a batch of a forward entry and a transposed entry of the same plan
    plan = pbb.Plan(ops, tgt, sources=src)
    wd = torch.tensor(w.reshape(ncols, ns), device="cuda")
    psi = torch.tensor(np.random.default_rng(92).standard_normal((ncols, nt)), device="cuda")
    o_f = torch.zeros((ncols, nt), dtype=torch.float64, device="cuda")
    o_t = torch.zeros((ncols, ns), dtype=torch.float64, device="cuda")
    batch = pbb.Batch([plan, plan], [wd, psi], [o_f, o_t], ncols, transpose=[False, True])
    for _ in range(2):
        batch.launch()
        torch.cuda.synchronize()
        assert torch.equal(o_f, plan.matvec(wd))
        assert torch.equal(o_t, plan.matvec_transpose(psi))


def test_separate_sets_reject_what_needs_one_set(gpu):
    import torch
    tgt, src = _sets(2000, 3000, 93)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 3, 4, 1e-5)
    plan = pbb.Plan(ops, tgt, sources=src)
    w = torch.zeros((1, 3000), dtype=torch.float64, device="cuda")
    with pytest.raises(pbb.capi.PBBError):
        plan.grad_positions(w, w)
    with pytest.raises(pbb.capi.PBBError):
        plan.update_points(np.vstack([tgt, src]))
    with pytest.raises(ValueError):
        plan.matvec(torch.zeros((1, 2000), dtype=torch.float64, device="cuda"))  # targets' count


def test_reject_outside(gpu):
    # points outside the root cube: extrapolated and counted by default, an
    # error with reject_outside
    xyz = np.concatenate([uniform_points(3000, seed=94) * 0.5, [[0.9, 0.9, 0.9]]])
    ops = pbb.Operators(pbb.laplacian(), 0.6, 3, 4, 1e-5)
    assert pbb.Plan(ops, xyz).info()["outside_points"] > 0
    with pytest.raises(pbb.capi.PBBError, match="outside the root cube"):
        pbb.Plan(ops, xyz, reject_outside=True)
    inside = pbb.Plan(ops, xyz[:-1], reject_outside=True)
    assert inside.info()["outside_points"] == 0


def test_plans_on_two_threads_and_streams(gpu):
    # two plans driven from two host threads, each on a stream of its own,
    # concurrently: every result equals the same matvec run alone
    import torch
    ops_a = pbb.Operators(pbb.laplacian(), 1.0, 4, 5, 1e-7)
    ops_b = pbb.Operators(pbb.gaussian(0.3), 1.0, 4, 4, 1e-6)
    pa = pbb.Plan(ops_a, clustered_points(20000, seed=95))
    tgt_b, src_b = _sets(9000, 12000, 96)
    pb = pbb.Plan(ops_b, tgt_b, sources=src_b)
    wa = torch.tensor(weights(20000, 4, seed=97).reshape(4, 20000), device="cuda")
    wb = torch.tensor(weights(12000, 3, seed=98).reshape(3, 12000), device="cuda")
    ref_a = pa.matvec(wa)
    ref_b = pb.matvec(wb, precision=pbb.capi.PREC_MIXED_F16X2)
    torch.cuda.synchronize()
    results, errors = {}, []

    def run(name, plan, w, prec):
        try:
            s = torch.cuda.Stream()
            outs = []
            with torch.cuda.stream(s):
                for _ in range(5):
                    outs.append(plan.matvec(w, precision=prec))
            s.synchronize()
            results[name] = outs
        except Exception as e:  # reported below, in the main thread
            errors.append(e)

    ts = [threading.Thread(target=run, args=("a", pa, wa, pbb.capi.PREC_FP64)),
          threading.Thread(target=run, args=("b", pb, wb, pbb.capi.PREC_MIXED_F16X2))]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    assert all(torch.equal(o, ref_a) for o in results["a"])
    assert all(torch.equal(o, ref_b) for o in results["b"])
