"""GMRES on the FMM operator, as the reference's Eigen application solves
(A + 1e4 I) x = b with its MatrixFMM and MatrixDirect (app/gmres.cpp:37-64):
on the reference's example input (input/*_test.bin: 5000 points, 1/r,
order 4, level 3, eps 1e-5), with scipy's GMRES over the plan-reusing
MatrixFMM, checked against the oracle's matvec, the direct-sum operator's
solution and a dense solve."""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc
from tests._helpers import oracle_ops_from
from tests.test_oracle import read_reference_input

pytestmark = pytest.mark.gpu


def test_gmres_on_reference_input(gpu):
    from scipy.sparse.linalg import gmres
    L, order, level, ncols, xyz, _ = read_reference_input()
    N = len(xyz)
    kernel = pbb.laplacian()
    ops = pbb.Operators(kernel, L, level, order, 1e-5)
    A = pbb.MatrixFMM(ops, xyz)
    D = pbb.MatrixDirect(kernel, xyz)
    b = np.random.default_rng(90).uniform(-1.0, 1.0, N)  # Eigen's setRandom
    # the operator is the oracle's FMM matvec plus 1e4 x
    x0 = np.random.default_rng(91).uniform(-1.0, 1.0, N)
    ref = orc.matvec(oracle_ops_from(ops, kernel), xyz, x0, 1) + 1e4 * x0
    assert orc.rel_err(A.matvec(x0), ref) <= 1e-10
    # the direct operator is the dense matrix
    diff = xyz[:, None, :] - xyz[None, :, :]
    r = np.sqrt((diff ** 2).sum(-1))
    with np.errstate(divide="ignore"):
        K = np.where(r > 0, 1.0 / r, 0.0)
    M = K + 1e4 * np.eye(N)
    assert orc.rel_err(D.matvec(x0), M @ x0) <= 1e-12
    xs = np.linalg.solve(M, b)
    xf, info_f = gmres(A.linear_operator(), b, rtol=1e-10, restart=30, maxiter=50)
    xd, info_d = gmres(D.linear_operator(), b, rtol=1e-10, restart=30, maxiter=50)
    assert info_f == 0 and info_d == 0
    print(f"GMRES products: FMM {A.products}, direct {D.products}; "
          f"FMM solution vs dense {orc.rel_err(xf, xs):.3e}, direct vs dense {orc.rel_err(xd, xs):.3e}")
    assert orc.rel_err(xd, xs) <= 1e-9
    # the FMM solves its own approximant: the solution is within the FMM's error of the exact one
    assert orc.rel_err(xf, xs) <= 1e-5
    assert orc.rel_err(A.matvec(xf), b) <= 1e-9
