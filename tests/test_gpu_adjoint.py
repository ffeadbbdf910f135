"""The adjoint (transposed FMM) through the C ABI.

Checks: the GPU transpose equals the transpose of the GPU forward operator
formed densely on a small problem (FP64, to rounding); <A x, y> = <x, A^T y>
at a realistic size; and parity with the oracle run on the transposed
operator set (U' = VT^T, VT' = U^T, E'_d = E_{315-d}^T) built in numpy.
"""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc
from tests._helpers import clustered_points, oracle_ops_from, uniform_points, weights

pytestmark = pytest.mark.gpu


def _t(a, n, dev="cuda:0"):
    import torch
    return torch.tensor(np.asarray(a).reshape(-1, n), dtype=torch.float64, device=dev)


def test_dense_transpose_small(gpu):
    import torch
    N = 300
    xyz = clustered_points(N, seed=31)
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 3, 4, 1e-7)
    plan = pbb.Plan(ops, xyz)
    I = torch.eye(N, dtype=torch.float64, device="cuda:0")  # N columns of unit weights
    A = plan.matvec(I).cpu().numpy()          # row c = A e_c  ->  A^T as an array
    At = plan.matvec_transpose(I).cpu().numpy()
    torch.cuda.synchronize()
    # A[c, i] = (A e_c)_i  => matrix M = A.T ; At[c, i] = (M^T e_c)_i => M^T = At.T
    M, MT = A.T, At.T
    assert np.linalg.norm(MT - M.T) <= 1e-13 * np.linalg.norm(M)


@pytest.mark.parametrize("prec,tol", [(pbb.capi.PREC_FP64, 1e-12), (pbb.capi.PREC_MIXED, 1e-6)],
                         ids=["fp64", "mixed"])
def test_inner_product_identity(gpu, prec, tol):
    import torch
    N, nc = 20000, 4
    xyz = uniform_points(N, seed=32)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 4, 5, 1e-7)
    plan = pbb.Plan(ops, xyz)
    x = _t(weights(N, nc, seed=33, dist="normal"), N)
    y = _t(weights(N, nc, seed=34, dist="normal"), N)
    Ax = plan.matvec(x, precision=prec)
    Aty = plan.matvec_transpose(y, precision=prec)
    torch.cuda.synchronize()
    lhs = float((Ax * y).sum())
    rhs = float((x * Aty).sum())
    assert abs(lhs - rhs) <= tol * float(Ax.norm() * y.norm())


def test_transpose_vs_oracle(gpu):
    import torch
    N, nc = 8000, 2
    xyz = clustered_points(N, seed=35)
    k = pbb.exponential(0.5)
    ops = pbb.Operators(k, 1.0, 4, 4, 1e-6)
    plan = pbb.Plan(ops, xyz)
    y = weights(N, nc, seed=36)
    got = plan.matvec_transpose(_t(y, N)).cpu().numpy().reshape(-1)
    torch.cuda.synchronize()
    o = oracle_ops_from(ops, k)
    n3, rf, rs = o.n ** 3, o.rf, o.rs
    Ut, VTt, Et = [], [], []
    for s in range(o.nsets):
        U = o.U[s * n3 * rf:(s + 1) * n3 * rf].reshape(rf, n3).T      # n3 x rf
        VT = o.VT[s * rs * n3:(s + 1) * rs * n3].reshape(n3, rs).T    # rs x n3
        E = o.E[s * 316 * rf * rs:(s + 1) * 316 * rf * rs].reshape(316, rs, rf).transpose(0, 2, 1)
        Ut.append(VT.T.ravel(order="F"))            # U' = VT^T  (n3 x rs)
        VTt.append(U.T.ravel(order="F"))            # VT' = U^T  (rf x n3)
        Et.append(np.stack([E[315 - d].T for d in range(316)]).transpose(0, 2, 1).ravel())
    adj = orc.Ops(o.kid, o.params, o.symmetry, o.homogen, o.L, o.levels, o.n, o.eps, rs, rf,
                  o.nsets, np.concatenate(Ut), np.concatenate(VTt), np.concatenate(Et))
    ref = orc.matvec(adj, xyz, y, nc)
    assert orc.rel_err(got, ref) <= 1e-10
