"""Shared helpers of the test suite (inputs, oracle runs, GPU runs)."""
import numpy as np

from oracle import oracle as orc


def uniform_points(n, seed=0):
    return np.random.default_rng(seed).random((n, 3))


def clustered_points(n, seed=0, nclusters=8, sigma=0.05, background=0.2):
    """Mixture of isotropic Gaussians (centres U[0.2,0.8]^3) plus a uniform
    background, clipped to the unit cube."""
    rng = np.random.default_rng(seed)
    nb = int(round(background * n))
    nc = n - nb
    centres = rng.uniform(0.2, 0.8, (nclusters, 3))
    lab = rng.integers(0, nclusters, nc)
    pts = centres[lab] + sigma * rng.standard_normal((nc, 3))
    pts = np.concatenate([pts, rng.random((nb, 3))])
    return np.clip(pts, 0.0, 1.0)


def weights(n, ncols, seed=1, dist="uniform"):
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        return rng.uniform(-1.0, 1.0, n * ncols)
    return rng.standard_normal(n * ncols)


def oracle_ops_from(ops, kernel):
    """The product's operators in the oracle's container."""
    U, VT, E, sf, ss = ops.export()
    return orc.Ops(kernel.id, np.array(list(kernel.params) + [0.0] * (4 - len(kernel.params))),
                   kernel.symmetry, kernel.homogen, ops.L, ops.tree_level, ops.order, ops.eps,
                   ops.rank_f, ops.rank_s, ops.nsets, U, VT, E, sf, ss)


def gpu_matvec(plan, w, ncols, device="cuda:0"):
    import torch
    N = plan.nsources  # the plan's points, or its separate sources
    wd = torch.tensor(np.asarray(w).reshape(ncols, N), dtype=torch.float64, device=device)
    out = plan.matvec(wd)
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(-1)


def gpu_direct(kernel, tgt, src, w, ncols, device="cuda:0"):
    """FP64 direct sum on the GPU (pbb_direct_sum, the DirectCalc3D of
    test.hpp:18-34): column-major (ncols * len(tgt),) numpy result for host
    inputs (w column-major ncols * len(src), or a CUDA tensor)."""
    import ctypes as C

    import torch

    from pbbfmm3d_amd import _capi as capi
    t = torch.tensor(np.ascontiguousarray(tgt), dtype=torch.float64, device=device)
    s = torch.tensor(np.ascontiguousarray(src), dtype=torch.float64, device=device)
    if isinstance(w, torch.Tensor):
        wd = w.to(device=device, dtype=torch.float64).contiguous()
    else:
        wd = torch.tensor(np.asarray(w).reshape(-1), dtype=torch.float64, device=device)
    out = torch.zeros(ncols * len(tgt), dtype=torch.float64, device=device)
    d = kernel.desc()
    capi.check(capi.lib().pbb_direct_sum(C.byref(d), t.data_ptr(), len(tgt), s.data_ptr(), len(src),
                                         wd.data_ptr(), ncols, out.data_ptr(), None))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def rows_of(idx, N, ncols):
    """Column-major row indices of points idx in every column."""
    idx = np.asarray(idx, dtype=np.int64)
    return (idx[None, :] + N * np.arange(ncols)[:, None]).ravel()
