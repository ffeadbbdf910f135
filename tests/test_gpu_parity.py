"""FP64 GPU matvec (through the C ABI) against the CPU oracle.

The oracle (oracle/pbb_oracle.c) restates compute.hpp; both sides use the
same operators (exported from the product) so that the comparison isolates
the matvec.  Bar: relative L2 difference <= 1e-10 (FP64, different summation
orders only).  Polynomial kernels of degree < n are reproduced exactly by
the Chebyshev FMM, so those cases are also checked against the direct sum.
"""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc
from tests._helpers import (clustered_points, gpu_matvec, oracle_ops_from, uniform_points,
                            weights)

PARITY = 1e-10

pytestmark = pytest.mark.gpu


def _run(kernel, L, level, n, eps, xyz, w, ncols):
    print(f"case kernel={kernel.id} L={L} level={level} n={n} eps={eps} N={len(xyz)} ncols={ncols}",
          flush=True)
    ops = pbb.Operators(kernel, L, level, n, eps)
    plan = pbb.Plan(ops, xyz)
    got = gpu_matvec(plan, w, ncols)
    ref = orc.matvec(oracle_ops_from(ops, kernel), xyz, w, ncols)
    return ops, plan, got, ref


def test_laplacian_20k_uniform_level4(gpu):
    # the reference's CPU-runnable case: 20k uniform, level 4, order 4, eps 1e-5
    xyz = uniform_points(20000, seed=0)
    w = weights(20000, 1, seed=1)
    ops, plan, got, ref = _run(pbb.laplacian(), 1.0, 4, 4, 1e-5, xyz, w, 1)
    assert ops.rank_f == 48 and ops.rank_s == 48
    assert orc.rel_err(got, ref) <= PARITY
    sample = np.arange(0, 20000, 97)
    d = orc.direct(orc.LAPLACIAN, None, xyz[sample], xyz, w, 1)
    assert orc.rel_err(got[sample], d) < 1e-3


def test_gaussian_per_level_operators(gpu):
    # non-homogeneous kernel: one operator set per level 2..4
    xyz = uniform_points(12000, seed=2)
    w = weights(12000, 3, seed=3)
    ops, plan, got, ref = _run(pbb.gaussian(0.3), 1.0, 4, 5, 1e-7, xyz, w, 3)
    assert ops.nsets == 3
    assert orc.rel_err(got, ref) <= PARITY


@pytest.mark.parametrize("kernel", [pbb.r2_homogeneous(), pbb.quadric(), pbb.constant()],
                         ids=["r2", "quadric", "constant"])
def test_polynomial_kernels_exact(gpu, kernel):
    xyz = clustered_points(5000, seed=4)
    w = weights(5000, 2, seed=5)
    ops, plan, got, ref = _run(kernel, 1.0, 4, 4, 1e-12, xyz, w, 2)
    assert orc.rel_err(got, ref) <= PARITY
    d = orc.direct(kernel.id, kernel.params, xyz, xyz, w, 2)
    assert orc.rel_err(got, d) <= 1e-12


def test_uniform_order5_ncols8(gpu):
    xyz = uniform_points(40000, seed=6)
    w = weights(40000, 8, seed=7)
    ops, plan, got, ref = _run(pbb.laplacian(), 1.0, 4, 5, 1e-7, xyz, w, 8)
    assert orc.rel_err(got, ref) <= PARITY


def test_clustered_wide_columns(gpu):
    # ncols > 64 exercises column chunking in M2L and the near field
    xyz = clustered_points(6000, seed=8)
    w = weights(6000, 70, seed=9, dist="normal")
    ops, plan, got, ref = _run(pbb.laplacian(), 1.0, 4, 4, 1e-5, xyz, w, 70)
    assert orc.rel_err(got, ref) <= PARITY


@pytest.mark.parametrize("level", [1, 2, 3, 5])
def test_levels(gpu, level):
    xyz = uniform_points(3000, seed=10 + level)
    w = weights(3000, 2, seed=11)
    ops, plan, got, ref = _run(pbb.laplacian(), 1.0, level, 4, 1e-6, xyz, w, 2)
    assert orc.rel_err(got, ref) <= PARITY


@pytest.mark.parametrize("n,kernel", [(3, pbb.gaussian(0.4)), (8, pbb.laplacian())], ids=["n3", "n8"])
def test_transfer_paths_by_order(gpu, n, kernel):
    # orders 4..8 run the direct per-level P2M / L2P passes and the split
    # near field; order 3 runs P2M + M2M + L2L and the one-kernel near
    # field; order 8 (n^3 = 512) also takes the generic GEMM
    xyz = clustered_points(4000, seed=50 + n)
    w = weights(4000, 3, seed=51)
    ops, plan, got, ref = _run(kernel, 1.0, 4, n, 1e-6, xyz, w, 3)
    assert orc.rel_err(got, ref) <= PARITY


def test_duplicates_and_tiny(gpu):
    # coincident points: 1/r at r = 0 is inf and is zeroed (compute.hpp:767)
    base = uniform_points(40, seed=12)
    xyz = np.concatenate([base, base[:10]])
    w = weights(len(xyz), 1, seed=13)
    ops, plan, got, ref = _run(pbb.laplacian(), 1.0, 3, 4, 1e-6, xyz, w, 1)
    assert np.all(np.isfinite(got))
    assert orc.rel_err(got, ref) <= PARITY
    d = orc.direct(orc.LAPLACIAN, None, xyz, xyz, w, 1)
    assert orc.rel_err(got, d) < 1e-3


def test_single_point(gpu):
    xyz = np.array([[0.3, 0.2, 0.1]])
    ops = pbb.Operators(pbb.gaussian(1.0), 1.0, 3, 4, 1e-6)
    plan = pbb.Plan(ops, xyz)
    got = gpu_matvec(plan, np.array([2.5]), 1)
    assert abs(got[0] - 2.5) < 1e-14  # K(x, x) = 1 for the Gaussian


def test_soft_laplacian_and_exponential(gpu):
    xyz = clustered_points(8000, seed=14)
    for kernel in (pbb.soft_laplacian(1e-6), pbb.exponential(0.5)):
        w = weights(8000, 4, seed=15, dist="normal")
        ops, plan, got, ref = _run(kernel, 1.0, 4, 5, 1e-7, xyz, w, 4)
        assert orc.rel_err(got, ref) <= PARITY


def test_output_accumulates_and_plan_reuse(gpu):
    import torch
    xyz = uniform_points(5000, seed=16)
    w = weights(5000, 2, seed=17)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 4, 4, 1e-5)
    plan = pbb.Plan(ops, xyz)
    wd = torch.tensor(w.reshape(2, 5000), dtype=torch.float64, device="cuda:0")
    out = torch.full_like(wd, 1.0)
    plan.matvec(wd, out)
    plan.matvec(wd, out)  # a plan serves any number of matvecs
    torch.cuda.synchronize()
    ref = orc.matvec(oracle_ops_from(ops, pbb.laplacian()), xyz, w, 2)
    got = out.cpu().numpy().reshape(-1)
    assert orc.rel_err(got - 1.0, 2 * ref) <= PARITY


def test_drop_in_host_path(gpu):
    # pbb_matvec: the H2_3D_Compute constructor's contract on host buffers
    xyz = uniform_points(6000, seed=18)
    w = weights(6000, 3, seed=19)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 4, 4, 1e-5)
    out = np.full(6000 * 3, 0.5)
    pbb.matvec_host(ops, xyz, w, 3, out)
    ref = orc.matvec(oracle_ops_from(ops, pbb.laplacian()), xyz, w, 3)
    assert orc.rel_err(out - 0.5, ref) <= PARITY


def test_result_ordered_on_torch_stream(gpu):
    # the matvec runs on torch's current stream (here the null stream):
    # reading the result back needs no device-wide synchronisation
    import torch
    xyz = uniform_points(3000, seed=40)
    w = weights(3000, 2, seed=41)
    ops = pbb.Operators(pbb.gaussian(0.5), 1.0, 3, 4, 1e-6)
    plan = pbb.Plan(ops, xyz)
    wd = torch.tensor(w.reshape(2, 3000), dtype=torch.float64, device="cuda:0")
    got = plan.matvec(wd).cpu().numpy().reshape(-1)
    ref = orc.matvec(oracle_ops_from(ops, pbb.gaussian(0.5)), xyz, w, 2)
    assert orc.rel_err(got, ref) <= PARITY


def test_operators_from_reference_cache_files(gpu, tmp_path):
    # operators read back from the reference-format cache give the same
    # matvec bit for bit
    xyz = clustered_points(6000, seed=30)
    w = weights(6000, 2, seed=31)
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 4, 5, 1e-7)
    ops.write_cache(str(tmp_path), "Gaussian")
    back = pbb.Operators.from_cache(pbb.gaussian(0.3), 1.0, 4, 5, 1e-7, str(tmp_path), "Gaussian")
    a = gpu_matvec(pbb.Plan(ops, xyz), w, 2)
    b = gpu_matvec(pbb.Plan(back, xyz), w, 2)
    assert np.array_equal(a, b)


def test_update_points_keeps_the_tree(gpu):
    # moving points inside their leaf boxes equals a plan built on the moved
    # points (same boxes, sorted order and centre) bit for bit, and a point
    # that leaves its leaf box is refused with the plan unchanged
    xyz = uniform_points(6000, seed=42)
    w = weights(6000, 2, seed=43)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 4, 4, 1e-5)
    plan = pbb.Plan(ops, xyz)
    before = gpu_matvec(plan, w, 2)
    ctr = 0.5 * (xyz.min(0) + xyz.max(0))
    rel = (xyz - ctr + 0.5) * 16  # leaf-box units at level 4
    inner = np.all(np.abs(rel - np.round(rel)) > 0.02, axis=1)
    inner[xyz.argmin(0)] = False
    inner[xyz.argmax(0)] = False
    moved = xyz.copy()
    moved[inner] += 1e-4  # < 0.02 leaf widths: every moved point stays in its box
    plan.update_points(moved)
    got = gpu_matvec(plan, w, 2)
    fresh = gpu_matvec(pbb.Plan(ops, moved), w, 2)
    assert np.array_equal(got, fresh)
    assert not np.array_equal(got, before)
    ref = orc.matvec(oracle_ops_from(ops, pbb.laplacian()), moved, w, 2)
    assert orc.rel_err(got, ref) <= PARITY
    out_of_box = moved.copy()
    k = int(np.flatnonzero(inner)[0])
    out_of_box[k, 0] += 1.0 / 16  # one leaf box over
    with pytest.raises(pbb.capi.PBBError):
        plan.update_points(out_of_box)
    assert np.array_equal(gpu_matvec(plan, w, 2), got)


@pytest.mark.parametrize("case", ["uniform-l4", "clustered-l7", "duplicates-l3", "one-point", "level1",
                                  "outside-l5"])
def test_gpu_tree_equals_host_tree(gpu, case):
    # the tree built on the GPU (pbb_plan_create) and the host restatement
    # (PBB_PLAN_HOST_TREE) give the same plan: same info, and matvecs equal
    # bit for bit (same point order, boxes, lists and work items)
    rng = np.random.default_rng(44)
    level, xyz = {
        "uniform-l4": (4, lambda: uniform_points(20000, seed=45)),
        "clustered-l7": (7, lambda: clustered_points(100000, seed=46)),
        "duplicates-l3": (3, lambda: np.repeat(uniform_points(500, seed=47), 4, axis=0)),
        "one-point": (3, lambda: np.array([[0.3, 0.4, 0.5]])),
        "level1": (1, lambda: uniform_points(700, seed=48)),
        "outside-l5": (5, lambda: np.concatenate([uniform_points(5000, seed=49) * 0.6,
                                                  [[0.0, 0.0, 0.0], [0.59, 0.59, 0.59]]])),
    }[case]
    xyz = xyz()
    L = 0.5 if case == "outside-l5" else 1.0  # points outside the root cube
    print(f"case {case} N={len(xyz)} level={level}", flush=True)
    ops = pbb.Operators(pbb.laplacian(), L, level, 4, 1e-5)
    g = pbb.Plan(ops, xyz)
    h = pbb.Plan(ops, xyz, host_tree=True)
    assert g.info() == h.info()
    w = rng.standard_normal(len(xyz) * 3)
    assert np.array_equal(gpu_matvec(g, w, 3), gpu_matvec(h, w, 3))


@pytest.mark.parametrize("kernel,n,eps", [(pbb.laplacian(), 7, 1e-12), (pbb.gaussian(0.3), 8, 1e-8)],
                         ids=["laplacian-n7-r313", "gaussian-n8-r250"])
def test_ranks_above_256_and_224(gpu, kernel, n, eps):
    # ranks above one workgroup's rows: the FP64 M2L splits them into row
    # groups above 256, the mixed M2L above 224 (row-wave tiles)
    xyz = clustered_points(6000, seed=70)
    w = weights(6000, 3, seed=71, dist="normal")
    ops, plan, got, ref = _run(kernel, 1.0, 4, n, eps, xyz, w, 3)
    print(f"ranks {ops.rank_f}/{ops.rank_s}", flush=True)
    assert ops.rank_f > 224
    assert orc.rel_err(got, ref) <= PARITY
    import torch
    wd = torch.tensor(w.reshape(3, -1), dtype=torch.float64, device="cuda:0")
    for prec in (pbb.capi.PREC_MIXED, pbb.capi.PREC_MIXED_F16X2):
        mix = plan.matvec(wd, precision=prec).cpu().numpy().reshape(-1)
        d = orc.rel_err(mix, got)
        print(f"precision {prec}: vs FP64 {d:.3e}", flush=True)
        assert d <= 1e-6


def test_column_chunks_when_the_workspace_is_capped(gpu):
    # a workspace cap below what 24 columns need runs the matvec in chunks of
    # columns: FP64 equals the one-pass matvec bit for bit, the mixed modes
    # stay within their bar of it
    import torch
    xyz = clustered_points(8000, seed=72)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 4, 4, 1e-5)
    plan = pbb.Plan(ops, xyz)
    w = torch.tensor(np.random.default_rng(73).standard_normal((24, 8000)), device="cuda")
    ref = plan.matvec(w)
    ref_m = plan.matvec(w, precision=pbb.capi.PREC_MIXED_F16X2)
    capped = pbb.Plan(ops, xyz)
    capped.set_workspace_limit(6 << 20)  # a few columns' worth
    got = capped.matvec(w)
    assert torch.equal(got, ref)
    got_m = capped.matvec(w, precision=pbb.capi.PREC_MIXED_F16X2)
    assert float(torch.linalg.norm(got_m - ref) / torch.linalg.norm(ref)) <= 1e-6
    assert float(torch.linalg.norm(got_m - ref_m) / torch.linalg.norm(ref)) <= 1e-6
