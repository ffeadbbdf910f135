"""Differentiation: the autograd function's gradients with respect to the
weights (transposed FMM) and to the point coordinates (differentiated
interpolants + kernel gradients), checked against central finite
differences of the FMM loss (h = 1e-4) and against the FP64 direct-sum
gradient."""
import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from tests._helpers import gpu_direct, uniform_points

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


def test_softened_laplacian_500k_level6(gpu):
    # 500k uniform points, level 6, order 5, 1/sqrt(r^2 + 1e-6), FP32 sigma
    # and b, 1 and 4 columns: the autograd gradients against central finite
    # differences (h = 1e-4, Richardson with h/2) of the FMM loss on 64
    # coordinates, and against the FP64 direct-sum gradient on 64 points
    import torch
    N, level, n, eps, h, delta = 500_000, 6, 5, 1e-7, 1e-4, 1e-6
    kernel = pbb.soft_laplacian(delta)
    xyz0 = uniform_points(N, seed=80)
    rng = np.random.default_rng(81)
    ops = pbb.Operators(kernel, 1.0, level, n, eps)
    plan = pbb.Plan(ops, xyz0)       # reused by the autograd calls
    fd_plan = pbb.Plan(ops, xyz0)    # moved by the finite differences
    picks = _safe_coords(xyz0, level, h, 64, rng, rmin=0.005)
    pts = rng.choice(N, 64, replace=False)
    for ncols in (1, 4):
        print(f"case N={N} ncols={ncols} level={level} order={n}", flush=True)
        sigma = torch.tensor(rng.standard_normal((ncols, N)), dtype=torch.float32, device="cuda")
        b = torch.tensor(rng.standard_normal((ncols, N)), dtype=torch.float32, device="cuda")
        xyz = torch.tensor(xyz0, dtype=torch.float64, requires_grad=True)
        sig = sigma.clone().requires_grad_(True)
        out = pbb.fmm_matvec(xyz, sig, ops, plan=plan)
        assert out.dtype == torch.float32
        L = 0.5 * ((out.double() - b.double()) ** 2).sum()
        L.backward()
        gx = xyz.grad.numpy()
        assert sig.grad.dtype == torch.float32
        s64, b64 = sigma.double(), b.double()

        def central(i, d, hh):
            o = []
            for sgn in (1, -1):
                xp = xyz0.copy()
                xp[i, d] += sgn * hh
                fd_plan.update_points(xp)
                o.append(fd_plan.matvec(s64))
            # L(+) - L(-) = 1/2 sum (o+ - o-)(o+ + o- - 2b), summed without
            # forming the two large losses
            return float(((o[0] - o[1]) * (o[0] + o[1] - 2 * b64)).sum()) / (4 * hh)

        errs = []
        for i, d in picks:
            fd = (4 * central(i, d, h / 2) - central(i, d, h)) / 3
            errs.append(abs(fd - gx[i, d]) / max(abs(fd), 1e-3 * np.abs(gx).max()))
        print(f"max relative FD mismatch over {len(picks)} coordinates: {max(errs):.3e}", flush=True)
        assert max(errs) < 1e-5
        fd_plan.update_points(xyz0)
        # exact gradients from FP64 direct sums: r = K sigma - b,
        # dL/dsigma = K r, dL/dx_i = sum_j f_ij (x_j - x_i) sum_c (r_ci s_cj + s_ci r_cj)
        # with f_ij = (|x_i - x_j|^2 + delta)^(-3/2)
        sn = s64.cpu().numpy()
        r = gpu_direct(kernel, xyz0, xyz0, sn.reshape(-1), ncols).reshape(ncols, N) - b64.cpu().numpy()
        gs_ref = gpu_direct(kernel, xyz0[pts], xyz0, r.reshape(-1), ncols).reshape(ncols, len(pts))
        gs = sig.grad.double().cpu().numpy()[:, pts]
        e_sig = np.linalg.norm(gs - gs_ref) / np.linalg.norm(gs_ref)
        gx_ref = np.zeros((len(pts), 3))
        for k, i in enumerate(pts):
            D = xyz0 - xyz0[i]
            f = ((D ** 2).sum(1) + delta) ** -1.5
            coef = r[:, i] @ sn + sn[:, i] @ r
            gx_ref[k] = (f * coef) @ D
        e_x = np.linalg.norm(gx[pts] - gx_ref) / np.linalg.norm(gx_ref)
        print(f"vs FP64 direct-sum gradients: sigma {e_sig:.3e}, positions {e_x:.3e}", flush=True)
        assert e_sig < 1e-4 and e_x < 1e-3
