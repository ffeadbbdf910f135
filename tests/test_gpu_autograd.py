"""Differentiation: the autograd function's gradients with respect to the
weights (transposed FMM) and to the point coordinates (differentiated
interpolants + kernel gradients), checked against central finite
differences of the FMM loss (h = 1e-4) and against the FP64 direct-sum
gradient."""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from tests._helpers import uniform_points

pytestmark = pytest.mark.gpu


def _loss(xyz, sigma, b, ops):
    import torch
    out = pbb.fmm_matvec(xyz, sigma, ops)
    return 0.5 * ((out - b) ** 2).sum()


def _safe_coords(xyz, level, h, k, rng, rmin=0.02):
    """k (point, axis) pairs >= 2h from every leaf-box face, not a bbox
    extreme (box membership and the centre stay fixed under +-h), and with
    no other point within rmin (the finite difference resolves the kernel)."""
    ctr = 0.5 * (xyz.min(0) + xyz.max(0))
    rel = (xyz - ctr + 0.5) * 2 ** level
    dist = np.abs(rel - np.round(rel)) / 2 ** level
    ok = dist > 2 * h
    from scipy.spatial import cKDTree
    dnn, _ = cKDTree(xyz).query(xyz, k=2)
    ok &= (dnn[:, 1] >= rmin)[:, None]
    for d in range(3):
        ok[np.argmin(xyz[:, d]), d] = False
        ok[np.argmax(xyz[:, d]), d] = False
    cand = np.argwhere(ok)
    return cand[rng.choice(len(cand), size=k, replace=False)]


@pytest.mark.parametrize("ncols", [1, 4])
def test_position_gradient_vs_finite_differences(gpu, ncols):
    import torch
    N, level, n, h = 6000, 4, 5, 1e-4
    xyz0 = uniform_points(N, seed=50)
    rng = np.random.default_rng(51)
    ops = pbb.Operators(pbb.soft_laplacian(1e-6), 1.0, level, n, 1e-7)
    sigma = torch.tensor(rng.standard_normal((ncols, N)), dtype=torch.float64, device="cuda")
    b = torch.tensor(rng.standard_normal((ncols, N)), dtype=torch.float64, device="cuda")
    xyz = torch.tensor(xyz0, dtype=torch.float64, requires_grad=True)
    sig = sigma.clone().requires_grad_(True)
    L = _loss(xyz, sig, b, ops)
    L.backward()
    gx = xyz.grad.numpy()
    # weights: dL/dsigma = A^T (A sigma - b)
    with torch.no_grad():
        plan = pbb.Plan(ops, xyz0)
        r = plan.matvec(sigma) - b
        gsig = plan.matvec_transpose(r)
    assert float((sig.grad - gsig).norm() / gsig.norm()) < 1e-12
    # positions: central differences of the FMM loss (h = 1e-4), Richardson-
    # extrapolated with h/2 so that the check resolves ~1e-8
    picks = _safe_coords(xyz0, level, h, 24, rng)

    def central(i, d, hh):
        vals = []
        for sgn in (1, -1):
            xp = xyz0.copy()
            xp[i, d] += sgn * hh
            with torch.no_grad():
                vals.append(float(_loss(torch.tensor(xp), sigma, b, ops)))
        return (vals[0] - vals[1]) / (2 * hh)

    errs = []
    for i, d in picks:
        fd = (4 * central(i, d, h / 2) - central(i, d, h)) / 3
        errs.append(abs(fd - gx[i, d]) / max(abs(fd), 1e-3 * np.abs(gx).max()))
    print(f"max relative FD mismatch over {len(picks)} coordinates: {max(errs):.3e}")
    assert max(errs) < 1e-5


def test_position_gradient_vs_direct_sum(gpu):
    import torch
    N, ncols = 3000, 2
    xyz0 = uniform_points(N, seed=52)
    rng = np.random.default_rng(53)
    sigma = rng.standard_normal((ncols, N))
    psi = rng.standard_normal((ncols, N))
    ops = pbb.Operators(pbb.soft_laplacian(1e-6), 1.0, 3, 6, 1e-9)
    plan = pbb.Plan(ops, xyz0)
    g = plan.grad_positions(torch.tensor(sigma, device="cuda"), torch.tensor(psi, device="cuda"))
    g = g.cpu().numpy()
    # exact gradient of sum_c psi_c^T K(x) sigma_c for K = (r^2 + 1e-6)^(-1/2)
    D = xyz0[None, :, :] - xyz0[:, None, :]          # D[i, j] = x_j - x_i
    r2 = (D ** 2).sum(-1)
    dK = (r2 + 1e-6) ** -1.5                          # grad_i K(x_i, x_j) = dK * (x_j - x_i)
    W = psi.T @ sigma                                 # W[i, j] = sum_c psi_ci sigma_cj
    Wsym = W + W.T
    ref = np.einsum("ij,ijd->id", dK * Wsym, D)
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    print(f"FMM vs direct-sum position gradient: {rel:.3e}")
    assert rel < 1e-4
