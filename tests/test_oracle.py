"""Pins the CPU oracle (oracle/pbb_oracle.c) before anything is checked
against it.  Each check is independent of the oracle's own formula: numpy
closed forms, brute force, or mathematical identities (interpolation,
partition of unity, polynomial reproduction, exactness of the FMM for
polynomial kernels of degree < n, which reduces the whole matvec to the
direct sum).
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests._helpers import clustered_points, uniform_points, weights

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def lagrange_cheb(n, x):
    """Independent S_m(x): the Lagrange basis on the Chebyshev nodes."""
    nodes = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    out = np.ones(n)
    for m in range(n):
        for j in range(n):
            if j != m:
                out[m] *= (x - nodes[j]) / (nodes[m] - nodes[j])
    return out


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8])
def test_closed_forms(n):
    nodes = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    np.testing.assert_allclose(orc.cheb_nodes(n), nodes, rtol=0, atol=1e-15)
    T = orc.tkz(n).reshape(n, n)
    np.testing.assert_allclose(T, np.cos(np.outer(np.arccos(nodes), np.arange(n))), atol=1e-13)
    w1 = 1 / np.sqrt(1 - nodes ** 2)
    kw = (w1[:, None, None] * w1[None, :, None] * w1[None, None, :]).ravel()
    np.testing.assert_allclose(orc.kweights(n), kw, rtol=1e-14)


@pytest.mark.parametrize("n", [3, 4, 6, 7])
def test_sn_identities(n):
    nodes = orc.cheb_nodes(n)
    for j in range(n):  # interpolation identity
        np.testing.assert_allclose(orc.sn(n, nodes[j]), np.eye(n)[j], atol=1e-13)
    rng = np.random.default_rng(n)
    for x in rng.uniform(-1.2, 1.2, 20):
        s = orc.sn(n, x)
        assert abs(s.sum() - 1.0) < 1e-12  # partition of unity
        np.testing.assert_allclose(s, lagrange_cheb(n, x), atol=1e-11)
        p = np.polynomial.Polynomial(rng.standard_normal(n))  # degree n-1
        assert abs(p(nodes) @ s - p(x)) < 1e-10 * (1 + abs(p(x)))


@pytest.mark.parametrize("n", [3, 4, 5])
def test_cweights(n):
    C = orc.cweights(n).reshape(2, n, n)
    nodes = orc.cheb_nodes(n)
    for half, shift in ((0, -0.5), (1, 0.5)):
        for j in range(n):
            np.testing.assert_allclose(C[half, :, j], lagrange_cheb(n, nodes[j] / 2 + shift),
                                       atol=1e-12)


def _p2m(n, u, w):
    """S[l1, l2, l3] = sum_j w_j S_l1(u_j0) S_l2(u_j1) S_l3(u_j2) (one column)."""
    S = np.zeros((n, n, n))
    for uj, wj in zip(u, w):
        sx, sy, sz = (orc.sn(n, c) for c in uj)
        S += wj * np.einsum("a,b,c->abc", sx, sy, sz)
    return S


@pytest.mark.parametrize("n", [4, 6])
def test_direct_passes_equal_m2m_and_l2l(n):
    """The GPU path's direct upward pass (P2M at every level) and downward
    pass (L2P from every level) rest on two identities of the reference's
    transfers (compute.hpp:956-1034 M2M, :798-878 L2L, Cweights
    H2_3D_Tree.cpp:805-819):
      M2M(P2M of the 8 children) == P2M at the parent, and
      L2P(L2L(F_parent)) at a child point == L2P(F_parent) at that point,
    because the parent's Chebyshev polynomials have degree n-1 and are
    interpolated exactly at the child's nodes."""
    rng = np.random.default_rng(n)
    C = orc.cweights(n).reshape(2, n, n)  # C[half][m][j] = S_m(x_j / 2 -+ 1/2)
    x = rng.uniform(-1, 1, (60, 3))       # points in the parent frame
    w = rng.standard_normal(60)
    Sp = np.zeros((n, n, n))
    Fp = rng.standard_normal((n, n, n))
    for o in range(8):
        bits = np.array([(o >> 2) & 1, (o >> 1) & 1, o & 1])
        sel = np.all((x >= 0) == bits.astype(bool), axis=1)
        centre = bits - 0.5
        uc = 2.0 * (x[sel] - centre)      # child frame
        Sc = _p2m(n, uc, w[sel])
        Cx, Cy, Cz = C[bits[0]], C[bits[1]], C[bits[2]]
        Sp += np.einsum("ai,bj,ck,ijk->abc", Cx, Cy, Cz, Sc)
        Fc = np.einsum("ai,bj,ck,abc->ijk", Cx, Cy, Cz, Fp)  # L2L = M2M^T
        for xj, ucj in zip(x[sel], uc):
            via_child = np.einsum("a,b,c,abc->", *(orc.sn(n, c) for c in ucj), Fc)
            direct = np.einsum("a,b,c,abc->", *(orc.sn(n, c) for c in xj), Fp)
            assert abs(via_child - direct) <= 1e-12 * (1 + np.abs(Fp).sum())
    np.testing.assert_allclose(Sp, _p2m(n, x, w), atol=1e-12 * np.abs(w).sum())


def test_ktable():
    K = orc.ktable()
    offs = [(a, b, c) for a in range(-3, 4) for b in range(-3, 4) for c in range(-3, 4)]
    far = [i for i, o in enumerate(offs) if max(map(abs, o)) > 1]
    assert len(far) == 316
    assert list(K[far]) == list(range(316))
    assert all(K[i] == -1 for i, o in enumerate(offs) if max(map(abs, o)) <= 1)


@pytest.mark.parametrize("kid,params", [(orc.LAPLACIAN, None), (orc.GAUSSIAN, [0.3]),
                                        (orc.EXPONENTIAL, [0.5, 1.0]),
                                        (orc.SOFT_LAPLACIAN, [1e-6]), (orc.LOGARITHM, None),
                                        (orc.THIN_PLATE, None), (orc.ONEOVERR2, None)])
def test_kernel_formulas(kid, params):
    rng = np.random.default_rng(kid)
    for _ in range(10):
        t, s = rng.random(3), rng.random(3)
        r = np.linalg.norm(s - t)
        expect = {orc.LAPLACIAN: 1 / r, orc.GAUSSIAN: np.exp(-r * r / 0.09),
                  orc.EXPONENTIAL: np.exp(-r / 0.5), orc.SOFT_LAPLACIAN: 1 / np.sqrt(r * r + 1e-6),
                  orc.LOGARITHM: np.log(r), orc.THIN_PLATE: r * r * np.log(r),
                  orc.ONEOVERR2: 1 / r ** 2}[kid]
        assert abs(orc.kernel_eval(kid, params, t, s) - expect) <= 1e-13 * abs(expect)


def test_kernel_blocks_brute_force_and_symmetry():
    n, L = 3, 0.7
    B = orc.kernel_blocks(orc.LAPLACIAN, None, 1, n, L)  # [d, j, i]
    Bfull = orc.kernel_blocks(orc.LAPLACIAN, None, 0, n, L)  # all 316 evaluated
    np.testing.assert_allclose(B, Bfull, rtol=1e-13)
    nodes = orc.cheb_nodes(n)
    kw = orc.kweights(n)
    offs = [(a, b, c) for a in range(-3, 4) for b in range(-3, 4) for c in range(-3, 4)
            if max(abs(a), abs(b), abs(c)) > 1]
    idx = [(i1, i2, i3) for i1 in range(n) for i2 in range(n) for i3 in range(n)]
    for d in (0, 100, 315):
        k = np.array(offs[d])
        for i in (0, 5, 26):
            for j in (1, 13):
                t = 0.5 * nodes[list(idx[i])] * L
                s = (k + 0.5 * nodes[list(idx[j])]) * L
                expect = 1 / np.linalg.norm(s - t) / (kw[i] * kw[j])
                assert abs(B[d, j, i] - expect) <= 1e-13 * expect


def test_cutoff_rule():
    rng = np.random.default_rng(3)
    for _ in range(20):
        s = np.sort(rng.lognormal(0, 4, 50))[::-1]
        eps = 10 ** rng.uniform(-9, -2)
        tail = np.cumsum((s ** 2)[::-1])[::-1]  # tail[k] = sum_{i>=k} s_i^2
        expect = int(np.sum(tail >= eps ** 2 * np.sum(s ** 2)))
        assert orc.cutoff(s, eps) == expect


def test_cutoff_rule_tie():
    # the reference keeps a value whose tail energy equals eps^2 * total
    # exactly (strict '<' drops, H2_3D_Tree.cpp:415-421): four unit values,
    # eps 1/2, tail of the last one = 1 = eps^2 * 4 -> all four kept
    assert orc.cutoff(np.ones(4), 0.5) == 4
    assert orc.cutoff(np.array([1.0, 1.0, 1.0, 0.5]), 0.5) == 3


def test_compress_identity():
    ops = orc.precompute(orc.LAPLACIAN, None, 1, -1.0, 1.0, 3, 3, 1e-4)
    n3 = 27
    blocks = orc.kernel_blocks(orc.LAPLACIAN, None, 1, 3, 1.0)
    U = ops.U.reshape(ops.rf, n3).T
    VT = ops.VT.reshape(n3, ops.rs).T
    E = ops.E.reshape(316, ops.rs, ops.rf).transpose(0, 2, 1)
    for d in (0, 157, 200):
        np.testing.assert_allclose(E[d], U.T @ blocks[d].T @ VT.T, atol=1e-12)
    # compressed operator approximates the weighted kernel to ~eps
    for d in (0, 200):
        K = blocks[d].T
        assert np.linalg.norm(U @ E[d] @ VT - K) <= 1e-3 * np.linalg.norm(K)


def test_direct_brute_force():
    xyz = uniform_points(300, seed=1)
    w = weights(300, 2, seed=2)
    out = orc.direct(orc.LAPLACIAN, None, xyz, xyz, w, 2)
    D = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=2)
    with np.errstate(divide="ignore"):
        K = np.where(D > 0, 1 / D, 0.0)
    ref = np.concatenate([K @ w[:300], K @ w[300:]])
    np.testing.assert_allclose(out, ref, rtol=1e-12)


@pytest.mark.parametrize("kid,homogen,n,levels", [(orc.R2, 2.0, 3, 3), (orc.R2, 2.0, 4, 4),
                                                  (orc.QUADRIC, 0.0, 4, 3),
                                                  (orc.CONSTANT, 0.0, 2, 4)])
def test_matvec_exact_for_polynomial_kernels(kid, homogen, n, levels):
    # Chebyshev interpolation of degree < n reproduces these kernels, so
    # the FMM equals the direct sum: every phase is pinned by this.
    xyz = clustered_points(2500, seed=levels)
    w = weights(2500, 3, seed=n)
    ops = orc.precompute(kid, None, 1, homogen, 1.0, levels, n, 1e-14)
    got = orc.matvec(ops, xyz, w, 3)
    ref = orc.direct(kid, None, xyz, xyz, w, 3)
    assert orc.rel_err(got, ref) < 1e-12


def test_matvec_level1_is_direct_sum():
    # one level: every box is a neighbour of every other -> all P2P
    xyz = uniform_points(700, seed=5)
    w = weights(700, 2, seed=6)
    ops = orc.precompute(orc.LAPLACIAN, None, 1, -1.0, 1.0, 1, 4, 1e-5)
    got = orc.matvec(ops, xyz, w, 2)
    ref = orc.direct(orc.LAPLACIAN, None, xyz, xyz, w, 2)
    assert orc.rel_err(got, ref) < 1e-13


def test_matvec_converges_with_order():
    xyz = uniform_points(3000, seed=7)
    w = weights(3000, 1, seed=8)
    ref = orc.direct(orc.LAPLACIAN, None, xyz, xyz, w, 1)
    errs = []
    for n in (3, 4, 5, 6):
        ops = orc.precompute(orc.LAPLACIAN, None, 1, -1.0, 1.0, 3, n, 1e-9)
        errs.append(orc.rel_err(orc.matvec(ops, xyz, w, 1), ref))
    assert errs[-1] < 1e-4
    assert all(b < a for a, b in zip(errs, errs[1:]))


def test_planted_mistake_is_caught():
    # a plausible bug -- one wrong far-field direction -- must fail the
    # polynomial-kernel exactness check
    xyz = uniform_points(2000, seed=9)
    w = weights(2000, 1, seed=10)
    ops = orc.precompute(orc.R2, None, 1, 2.0, 1.0, 3, 3, 1e-14)
    E = ops.E.reshape(316, -1).copy()
    E[[10, 11]] = E[[11, 10]]  # swap two directions' operators
    bad = orc.Ops(**{**ops.__dict__, "E": E.reshape(-1)})
    ref = orc.direct(orc.R2, None, xyz, xyz, w, 1)
    assert orc.rel_err(orc.matvec(ops, xyz, w, 1), ref) < 1e-12
    assert orc.rel_err(orc.matvec(bad, xyz, w, 1), ref) > 1e-6


def read_reference_input():
    d = os.path.join(GOLDEN, "reference_input")
    meta = open(os.path.join(d, "metadata_test.txt")).read().strip().split(",")
    L, order, Ns, Nf, ncols, level = float(meta[0]), int(meta[1]), int(meta[2]), int(meta[3]), \
        int(meta[4]), int(meta[5])
    f = np.fromfile(os.path.join(d, "field_test.bin"), dtype=np.float64)
    xyz = f[: 3 * Nf].reshape(3, Nf).T.copy()  # x block, y block, z block
    q = np.fromfile(os.path.join(d, "charge_test.bin"), dtype=np.float64)[: Ns * ncols]
    return L, order, level, ncols, xyz, q


def test_reference_input_fixture():
    # the reference's own example input (input/*_test.bin, read the way
    # read_sources.cpp:20-60 reads it), Laplacian, order 4, level 3, eps 1e-5
    # as binary_file_standard_kernel.cpp:18-55 runs it
    L, order, level, ncols, xyz, q = read_reference_input()
    assert xyz.shape == (5000, 3) and np.all(np.abs(xyz) <= 0.5)
    ops = orc.precompute(orc.LAPLACIAN, None, 1, -1.0, L, level, order, 1e-5)
    got = orc.matvec(ops, xyz, q, ncols)
    ref = orc.direct(orc.LAPLACIAN, None, xyz, xyz, q, ncols)
    assert orc.rel_err(got, ref) < 1e-4
    golden = np.load(os.path.join(GOLDEN, "reference_input_direct.npy"))
    np.testing.assert_allclose(ref, golden, rtol=1e-12)


@pytest.mark.parametrize("kid,homogen,params,levels,n,pts", [
    (orc.LAPLACIAN, -1.0, None, 4, 4, "uniform"),
    (orc.GAUSSIAN, 0.0, [0.3], 4, 4, "clustered"),
    (orc.LAPLACIAN, -1.0, None, 5, 3, "clustered"),
])
def test_masked_matvec_equals_full(kid, homogen, params, levels, n, pts):
    # the target-restricted oracle forms the same outputs on its targets
    N, ncols = 6000, 2
    xyz = uniform_points(N, seed=21) if pts == "uniform" else clustered_points(N, seed=21)
    w = weights(N, ncols, seed=22)
    ops = orc.precompute(kid, params, 1, homogen, 1.0, levels, n, 1e-6)
    full = orc.matvec(ops, xyz, w, ncols)
    tgt = np.random.default_rng(23).choice(N, 150, replace=False)
    sentinel = np.full(N * ncols, 7.0)
    part = orc.matvec(ops, xyz, w, ncols, out=sentinel.copy(), targets=tgt)
    rows = (tgt[None, :] + N * np.arange(ncols)[:, None]).ravel()
    got = part[rows] - 7.0
    assert orc.rel_err(got, full[rows]) < 1e-13
    # outputs of points outside the targets' leaves are left as they are
    assert np.count_nonzero(part == 7.0) > N * ncols // 2


@pytest.mark.parametrize("kid,homogen,n,levels", [(orc.R2, 2.0, 3, 3), (orc.QUADRIC, 0.0, 4, 4)])
def test_matvec_ts_exact_for_polynomial_kernels(kid, homogen, n, levels):
    # separate targets and sources (different counts and bboxes; a few
    # sources outside the targets' root cube): for a polynomial kernel of
    # degree < n the FMM is exact, so it must equal the direct sum
    tgt = clustered_points(1800, seed=31)
    src = 0.9 * uniform_points(2600, seed=32) + 0.04
    w = weights(2600, 2, seed=33)
    ops = orc.precompute(kid, None, 1, homogen, 1.0, levels, n, 1e-14)
    got = orc.matvec_ts(ops, tgt, src, w, 2)
    ref = orc.direct(kid, None, tgt, src, w, 2)
    assert got.shape == (1800 * 2,)
    assert orc.rel_err(got, ref) < 1e-12


def test_matvec_ts_converges_to_direct_sum():
    tgt = uniform_points(1500, seed=34)
    src = clustered_points(2500, seed=35)
    w = weights(2500, 1, seed=36)
    ref = orc.direct(orc.GAUSSIAN, [0.3], tgt, src, w, 1)
    errs = []
    for n in (4, 6):
        ops = orc.precompute(orc.GAUSSIAN, [0.3], 1, 0.0, 1.0, 3, n, 1e-10)
        errs.append(orc.rel_err(orc.matvec_ts(ops, tgt, src, w, 1), ref))
    assert errs[1] < errs[0] and errs[1] < 1e-5


def test_matvec_ts_of_the_same_points_is_matvec():
    # sources equal to the targets: every point appears twice (once without
    # weight), the sums are the same terms
    xyz = clustered_points(3000, seed=37)
    w = weights(3000, 2, seed=38)
    ops = orc.precompute(orc.LAPLACIAN, None, 1, -1.0, 1.0, 4, 4, 1e-6)
    assert orc.rel_err(orc.matvec_ts(ops, xyz, xyz, w, 2), orc.matvec(ops, xyz, w, 2)) < 1e-13
