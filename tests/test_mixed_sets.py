"""The mixed modes' re-compressed operator sets (pbb_ops_mixed_ranks), on
the host: the ranks kept per set against numpy's SVD of the same stacked
M2L blocks, and the accuracy the kept directions carry.

The check is independent of the product's TSQR + Jacobi SVD: numpy/LAPACK
singular values of [E_0 .. E_315] and [E_0; ..; E_315] give the smallest
kf (multiple of 16) and ks (16 or a multiple of 32) whose discarded values
hold at most tol of the Frobenius norm; the projection of every E_d on those
leading directions must reproduce E_d within that bound.
"""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb


def _stacks(ops):
    U, VT, E, sf, ss = ops.export()
    rf, rs = ops.rank_f, ops.rank_s
    E = E.reshape(ops.nsets, 316, rs, rf).transpose(0, 1, 3, 2)  # [set][d] rf x rs
    return E


def _rank(sigma, r, first, mult, tol):
    tot = np.sum(sigma ** 2)
    rest = np.concatenate([np.cumsum((sigma ** 2)[::-1])[::-1], [0.0]])
    for k in [first] + list(range(mult, r, mult)):
        if k < r and rest[k] <= tol * tol * tot:
            return k
    return r


@pytest.mark.parametrize("tol", [1e-10, 1e-8, 2e-8])
def test_ranks_against_numpy(tol):
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 5, 5, 1e-8, host_precompute=True)
    ops.set_mixed_tolerance(tol)
    got = ops.mixed_ranks()
    E = _stacks(ops)
    rf, rs = ops.rank_f, ops.rank_s
    assert len(got) == ops.nsets == 4
    for s, (kf, ks, tail) in enumerate(got):
        A = np.concatenate(list(E[s]), axis=1)  # rf x 316 rs
        B = np.concatenate(list(E[s]), axis=0)  # 316 rf x rs
        UL, sL, _ = np.linalg.svd(A, full_matrices=False)
        _, sR, VRt = np.linalg.svd(B, full_matrices=False)
        kf_ref, ks_ref = _rank(sL, rf, 16, 16, tol), _rank(sR, rs, 16, 32, tol)
        if kf_ref >= rf and ks_ref >= rs:  # no gain: the set's own factors
            assert (kf, ks) == (rf, rs) and tail == 0.0
            continue
        # the boundary value itself may fall on either side of the bound
        # within rounding: allow one step
        assert abs(kf - kf_ref) <= 16 and abs(ks - ks_ref) <= 32, (s, kf, kf_ref, ks, ks_ref)
        assert ks == 16 or ks % 32 == 0 or ks == rs
        assert 0 < tail <= tol * 1.0001
        # the kept directions reproduce every block within the bound
        QL, QR = UL[:, :kf], VRt[:ks].T
        num = sum(np.linalg.norm(Ed - QL @ (QL.T @ Ed @ QR) @ QR.T) ** 2 for Ed in E[s])
        assert np.sqrt(num) <= 2.0 * tol * np.linalg.norm(A)
    # smooth kernel, small boxes: the deepest set keeps far fewer directions
    assert got[-1][0] < rf and got[-1][1] < rs


def test_homogeneous_kernel_keeps_its_rank_and_tolerance_rules():
    ops = pbb.Operators(pbb.laplacian(), 1.0, 4, 4, 1e-5, host_precompute=True)
    (kf, ks, tail), = ops.mixed_ranks()
    assert kf == ops.rank_f and ks == ops.rank_s and tail == 0.0
    # the tolerance is fixed once the sets are made
    with pytest.raises(pbb.capi.PBBError):
        ops.set_mixed_tolerance(1e-9)
    g = pbb.Operators(pbb.gaussian(0.3), 1.0, 4, 4, 1e-7, host_precompute=True)
    with pytest.raises(pbb.capi.PBBError):
        g.set_mixed_tolerance(1e-3)  # above the mixed modes' own bar
    g.set_mixed_tolerance(0.0)
    assert all(kf == g.rank_f and tail == 0.0 for kf, _, tail in g.mixed_ranks())
