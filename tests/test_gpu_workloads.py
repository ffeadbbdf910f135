"""The benchmark workloads at their full sizes, through the C ABI.

* 1,000,000 uniform points, tree level 6, order 5, eps 1e-7, 8 columns, 1/r
  (the drop-in configuration): the FP64 plan matvec and the one-shot host
  drop-in (pbb_matvec, as H2_3D_Compute) against the oracle on 2000 sampled
  targets (<= 1e-10), and against the FP64 direct sum there (the FMM's own
  approximation level).
* 200,000 clustered points at tree level 7, order 4: the deepest trees
  (level-7 coordinate maps, tile counts, one operator set per level for the
  Gaussian) against the oracle (<= 1e-10).
* 4,000,000 clustered points, level 7, order 6, eps 1e-8, 32 columns,
  exp(-r^2/0.3^2) and 1/r (the benchmark): the FP64 path against the FP64
  direct sum on 10,000 sampled targets (approximation level), and the f16x2
  path, with the M2L tile the accuracy-bounded autotuner keeps, within 1e-6
  of the FP64 path, on the tuning weights and on other weights.

The oracle runs target-restricted (oracle.matvec(..., targets=...)): it forms
the same outputs on the sampled targets as the full oracle (pinned in
tests/test_oracle.py), at a cost the host cores finish in seconds.
"""
import gc
import time

import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc
from tests._helpers import (clustered_points, gpu_direct, gpu_matvec, oracle_ops_from, rows_of,
                            uniform_points, weights)

pytestmark = pytest.mark.gpu

PARITY = 1e-10


def _free():
    import torch
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


def test_uniform_1m_level6_vs_oracle(gpu):
    N, nc = 1_000_000, 8
    kernel = pbb.laplacian()
    print(f"case 1/r N={N} ncols={nc} level=6 order=5 eps=1e-7", flush=True)
    xyz = uniform_points(N, seed=0)
    w = weights(N, nc, seed=1)
    ops = pbb.Operators(kernel, 1.0, 6, 5, 1e-7)
    assert (ops.rank_f, ops.rank_s) == (101, 101)
    plan = pbb.Plan(ops, xyz)
    got = gpu_matvec(plan, w, nc)
    t0 = time.perf_counter()
    host = pbb.matvec_host(ops, xyz, w, nc)
    print(f"drop-in pbb_matvec (plan + matvec + copies): {time.perf_counter() - t0:.2f} s", flush=True)
    tgt = np.sort(np.random.default_rng(2).choice(N, 2000, replace=False))
    rows = rows_of(tgt, N, nc)
    t0 = time.perf_counter()
    ref = orc.matvec(oracle_ops_from(ops, kernel), xyz, w, nc, targets=tgt)[rows]
    print(f"oracle on {len(tgt)} targets: {time.perf_counter() - t0:.1f} s "
          f"({orc.num_threads()} threads)", flush=True)
    e_plan, e_host = orc.rel_err(got[rows], ref), orc.rel_err(host[rows], ref)
    print(f"plan vs oracle {e_plan:.3e}, drop-in vs oracle {e_host:.3e}", flush=True)
    assert e_plan <= PARITY and e_host <= PARITY
    assert orc.rel_err(host, got) <= 1e-14
    d = gpu_direct(kernel, xyz[tgt], xyz, w, nc)
    e_direct = orc.rel_err(got[rows], d)
    print(f"FMM vs FP64 direct sum on the targets: {e_direct:.3e}", flush=True)
    assert e_direct <= 1e-4
    del plan
    _free()


@pytest.mark.parametrize("kname", ["laplacian", "gaussian"])
def test_clustered_200k_level7_vs_oracle(gpu, kname):
    N, nc = 200_000, 2
    kernel = {"laplacian": pbb.laplacian(), "gaussian": pbb.gaussian(0.3)}[kname]
    print(f"case {kname} N={N} ncols={nc} level=7 order=4 eps=1e-6", flush=True)
    xyz = clustered_points(N, seed=5)
    w = weights(N, nc, seed=6, dist="normal")
    ops = pbb.Operators(kernel, 1.0, 7, 4, 1e-6)
    assert ops.nsets == (1 if kname == "laplacian" else 6)
    plan = pbb.Plan(ops, xyz)
    info = plan.info()
    assert info["tree_level"] == 7 and info["boxes"][7] > 10_000
    got = gpu_matvec(plan, w, nc)
    tgt = np.sort(np.random.default_rng(7).choice(N, 3000, replace=False))
    rows = rows_of(tgt, N, nc)
    ref = orc.matvec(oracle_ops_from(ops, kernel), xyz, w, nc, targets=tgt)[rows]
    err = orc.rel_err(got[rows], ref)
    print(f"level-7 tree vs oracle: {err:.3e} (leaves {info['boxes'][7]}, m2l pairs "
          f"{info['m2l_pairs']})", flush=True)
    assert err <= PARITY
    del plan
    _free()


@pytest.mark.parametrize("kname", ["gaussian", "laplacian"])
def test_clustered_4m_level7_benchmark(gpu, kname):
    import torch
    N, nc = 4_000_000, 32
    kernel = {"laplacian": pbb.laplacian(), "gaussian": pbb.gaussian(0.3)}[kname]
    f16x2 = pbb.capi.PREC_MIXED_F16X2
    print(f"case {kname} N={N} ncols={nc} level=7 order=6 eps=1e-8", flush=True)
    xyz = clustered_points(N, seed=0)
    w = torch.tensor(np.random.default_rng(1).standard_normal((nc, N)), device=gpu)
    ops = pbb.Operators(kernel, 1.0, 7, 6, 1e-8)
    plan = pbb.Plan(ops, xyz)
    out64 = plan.matvec(w)
    ts = np.sort(np.random.default_rng(2).choice(N, 10_000, replace=False))
    tsd = torch.tensor(ts, device=gpu)
    ref = torch.tensor(gpu_direct(kernel, xyz[ts], xyz, w, nc).reshape(nc, len(ts)), device=gpu)

    def rel(a, b):
        return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))

    e64 = rel(out64[:, tsd], ref)
    print(f"FP64 path vs FP64 direct sum on {len(ts)} targets: {e64:.3e}", flush=True)
    assert e64 <= 1e-5  # order-6 truncation (measured 5.8e-6 / 7.1e-6)
    tile = plan.tune_m2l(w, f16x2, max_rel_diff=1e-6)
    print(f"tile kept by the bounded tuner: {tile}", flush=True)
    assert tile["rel_diff"] <= 1e-6
    outm = plan.matvec(w, precision=f16x2)
    d_same = rel(outm, out64)
    em = rel(outm[:, tsd], ref)
    print(f"f16x2 vs FP64 path {d_same:.3e}, vs direct sum {em:.3e}", flush=True)
    assert d_same <= 1e-6 and em <= 1e-5
    # other weights of the same distribution
    w2 = torch.tensor(np.random.default_rng(3).standard_normal((nc, N)), device=gpu)
    d_other = rel(plan.matvec(w2, precision=f16x2), plan.matvec(w2))
    print(f"f16x2 vs FP64 path on other weights: {d_other:.3e}", flush=True)
    assert d_other <= 1e-6
    del plan, w, w2, out64, outm
    _free()
