"""The product's host FP64 operator precompute (own TSQR + Jacobi SVD),
called through the C ABI, against the oracle's numpy/LAPACK precompute.

SVD factors are unique only up to signs and, inside a cluster of equal
singular values, up to rotation; the comparisons use what is invariant:
singular values, cutoffs, orthonormality, U E V^T, and the projector U U^T
wherever the cut does not split a cluster.
"""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from oracle import oracle as orc

CASES = [  # kernel, L, levels, n, eps
    (pbb.laplacian(), 1.0, 4, 4, 1e-5),
    (pbb.laplacian(), 1.0, 4, 5, 1e-7),
    (pbb.gaussian(0.3), 1.0, 4, 4, 1e-7),
    (pbb.exponential(0.5), 2.0, 3, 4, 1e-6),
    (pbb.oneoverr2(), 1.0, 3, 3, 1e-4),
]


def _pair(kernel, L, levels, n, eps):
    ops = pbb.Operators(kernel, L, levels, n, eps)
    ref = orc.precompute(kernel.id, kernel.params, kernel.symmetry, kernel.homogen, L, levels, n, eps)
    return ops, ref


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"k{c[0].id}-n{c[3]}-eps{c[4]}")
def test_singular_values_and_cutoffs(case):
    ops, ref = _pair(*case)
    U, VT, E, sf, ss = ops.export()
    np.testing.assert_allclose(sf, ref.sigma_f, rtol=0, atol=1e-12 * ref.sigma_f[0])
    np.testing.assert_allclose(ss, ref.sigma_s, rtol=0, atol=1e-12 * ref.sigma_s[0])
    assert (ops.rank_f, ops.rank_s) == (ref.rf, ref.rs)
    assert ops.nsets == ref.nsets


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"k{c[0].id}-n{c[3]}-eps{c[4]}")
def test_operator_invariants(case):
    ops, ref = _pair(*case)
    n3, rf, rs = ops.n3, ops.rank_f, ops.rank_s
    U, VT, E, sf, _ = ops.export()
    kernel, L, levels, n, eps = case
    for s in range(ops.nsets):
        Us = U[s * n3 * rf:(s + 1) * n3 * rf].reshape(rf, n3).T
        Vs = VT[s * rs * n3:(s + 1) * rs * n3].reshape(n3, rs)  # V = VT^T
        Es = E[s * 316 * rf * rs:(s + 1) * 316 * rf * rs].reshape(316, rs, rf).transpose(0, 2, 1)
        np.testing.assert_allclose(Us.T @ Us, np.eye(rf), atol=1e-13)
        np.testing.assert_allclose(Vs.T @ Vs, np.eye(rs), atol=1e-13)
        box = 1.0 if ops.homogeneous else L / 4 / 2 ** s
        blocks = orc.kernel_blocks(kernel.id, kernel.params, kernel.symmetry, n, box)
        Ro = ref.U[s * n3 * rf:(s + 1) * n3 * rf].reshape(rf, n3).T
        Vo = ref.VT[s * rs * n3:(s + 1) * rs * n3].reshape(n3, rs)
        Eo = ref.E[s * 316 * rf * rs:(s + 1) * 316 * rf * rs].reshape(316, rs, rf).transpose(0, 2, 1)
        for d in (0, 77, 158, 315):
            K = blocks[d].T
            np.testing.assert_allclose(Es[d], Us.T @ K @ Vs, atol=1e-12 * np.abs(K).max())
            # both compressions approximate K to the truncation level, and
            # agree with each other to that level
            A = Us @ Es[d] @ Vs.T
            B = Ro @ Eo[d] @ Vo.T
            assert np.linalg.norm(A - K) <= 50 * eps * np.linalg.norm(K) + 1e-13
            assert np.linalg.norm(A - B) <= 50 * eps * np.linalg.norm(K) + 1e-13
        # projectors agree when the cut does not split a cluster
        gap = (sf[rf - 1] - sf[rf]) / sf[0] if rf < n3 else 1.0
        if gap > 1e-6:
            np.testing.assert_allclose(Us @ Us.T, Ro @ Ro.T, atol=1e-8)


def test_host_callback_matches_functor():
    # a host callback (the virtual EvaluateKernel path) gives the same ops
    k = pbb.gaussian(0.3)
    a = pbb.Operators(k, 1.0, 3, 4, 1e-6)
    b = pbb.Operators(k, 1.0, 3, 4, 1e-6, host_eval=lambda t, s: k.eval(t, s))
    Ua, VTa, Ea, sfa, ssa = a.export()
    Ub, VTb, Eb, sfb, ssb = b.export()
    np.testing.assert_allclose(sfa, sfb, rtol=0, atol=1e-13 * sfa[0])
    np.testing.assert_allclose(ssa, ssb, rtol=0, atol=1e-13 * ssa[0])
    assert (a.rank_f, a.rank_s) == (b.rank_f, b.rank_s)
    n3, rf, rs = a.n3, a.rank_f, a.rank_s
    for s in range(a.nsets):  # U E V^T is invariant under the SVD's freedom
        def full(U, VT, E, d):
            Us = U[s * n3 * rf:(s + 1) * n3 * rf].reshape(rf, n3).T
            Vs = VT[s * rs * n3:(s + 1) * rs * n3].reshape(n3, rs)
            Es = E[(s * 316 + d) * rf * rs:(s * 316 + d + 1) * rf * rs].reshape(rs, rf).T
            return Us @ Es @ Vs.T
        for d in (0, 200):
            A, B = full(Ua, VTa, Ea, d), full(Ub, VTb, Eb, d)
            assert np.linalg.norm(A - B) <= 1e-5 * np.linalg.norm(A)


def test_from_arrays_roundtrip():
    a = pbb.Operators(pbb.laplacian(), 1.0, 3, 4, 1e-5)
    U, VT, E, _, _ = a.export()
    b = pbb.Operators.from_arrays(pbb.laplacian(), 1.0, 3, 4, 1e-5, a.rank_f, a.rank_s, a.nsets,
                                  U, VT, E)
    for x, y in zip(a.export()[:3], b.export()[:3]):
        np.testing.assert_array_equal(x, y)


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,n,level", [(pbb.laplacian(), 5, 3), (pbb.gaussian(0.3), 6, 4)],
                         ids=["laplacian-n5", "gaussian-n6"])
def test_gpu_precompute_equals_host(gpu, kernel, n, level):
    # kernel blocks and the E compression on the GPU, SVDs on the host: the
    # same singular values, cutoffs and compressed operators U E V^T as the
    # all-host precompute, and the same FMM to the SVD truncation
    import torch
    eps = 1e-7
    g = pbb.Operators(kernel, 1.0, level, n, eps)
    h = pbb.Operators(kernel, 1.0, level, n, eps, host_precompute=True)
    Ug, VTg, Eg, sfg, ssg = g.export()
    Uh, VTh, Eh, sfh, ssh = h.export()
    assert (g.rank_f, g.rank_s) == (h.rank_f, h.rank_s)
    assert np.max(np.abs(sfg - sfh)) <= 1e-12 * sfh[0] and np.max(np.abs(ssg - ssh)) <= 1e-12 * ssh[0]
    n3, rf, rs = g.n3, g.rank_f, g.rank_s
    for s in range(g.nsets):
        def full(U, VT, E, d):
            Us = U[s * n3 * rf:(s + 1) * n3 * rf].reshape(rf, n3).T
            Vs = VT[s * rs * n3:(s + 1) * rs * n3].reshape(n3, rs).T
            Ed = E[(s * 316 + d) * rf * rs:(s * 316 + d + 1) * rf * rs].reshape(rs, rf).T
            return Us @ Ed @ Vs
        # the truncation error is relative to the whole operator set, so the
        # blocks are compared against the largest block's norm
        pairs = [(full(Ug, VTg, Eg, d), full(Uh, VTh, Eh, d)) for d in range(0, 316, 15)]
        top = max(np.linalg.norm(b) for _, b in pairs)
        assert max(np.linalg.norm(a - b) for a, b in pairs) <= 10 * eps * top
    xyz = np.random.default_rng(5).random((6000, 3))
    w = torch.tensor(np.random.default_rng(6).standard_normal((2, 6000)), device="cuda")
    yg = pbb.Plan(g, xyz).matvec(w)
    yh = pbb.Plan(h, xyz).matvec(w)
    assert float(torch.linalg.norm(yg - yh) / torch.linalg.norm(yh)) <= 10 * eps


def test_deeper_levels_keep_the_level2_cutoff():
    # A non-homogeneous kernel gets one operator set per level, and every set
    # keeps the cutoff of the level-2 blocks (H2_3D_Tree.cpp:406-436), although
    # the blocks of smaller boxes have a lower numerical rank.  Pinned against
    # numpy SVDs of each level's stacked kernel blocks: the ranks are the
    # level-2 cutoffs, and the deepest level's own cutoff is smaller, so a
    # per-level cutoff would fail here.
    # (cutoffs of levels 2..5 here: 54, 32, 18, 10 of n^3 = 64)
    kernel, L, levels, n, eps = pbb.gaussian(0.5), 1.0, 5, 4, 1e-6
    ops = pbb.Operators(kernel, L, levels, n, eps)
    U, VT, E, sf, ss = ops.export()
    own = []
    for s in range(ops.nsets):
        B = orc.kernel_blocks(kernel.id, kernel.params, kernel.symmetry, n, L / 4 / 2 ** s)
        fat = np.hstack([B[d].T for d in range(316)])
        thin = np.vstack([B[d].T for d in range(316)])
        sv_f = np.linalg.svd(fat, compute_uv=False)
        sv_s = np.linalg.svd(thin, compute_uv=False)
        if s == 0:
            k = min(len(sf), len(sv_f))
            np.testing.assert_allclose(sf[:k], sv_f[:k], rtol=0, atol=1e-12 * sv_f[0])
            np.testing.assert_allclose(ss[:k], sv_s[:k], rtol=0, atol=1e-12 * sv_s[0])
        own.append((orc.cutoff(sv_f, eps), orc.cutoff(sv_s, eps)))
    assert ops.nsets == levels - 1
    assert (ops.rank_f, ops.rank_s) == own[0] and ops.rank_f < ops.n3
    assert own[-1][0] < own[0][0] and own[-1][1] < own[0][1]
    rf, rs, n3 = ops.rank_f, ops.rank_s, ops.n3
    assert U.size == ops.nsets * n3 * rf and VT.size == ops.nsets * rs * n3
    assert E.size == ops.nsets * 316 * rf * rs
