"""pbb_plan_matvec_noalloc (the matvec whose workspace is the library's fixed
per-GPU array) against pbb_plan_matvec, through the C ABI.

Two calls in a row, the second on a smaller plan with fewer columns, so the
second call runs in an array that still holds the first call's data; a
mixed-precision call after its operators were split; and a matvec too large
for the array, which must fail before any launch."""
import ctypes as C

import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from pbbfmm3d_amd import _capi as capi

pytestmark = pytest.mark.gpu


def _noalloc(plan, w, out, ncols, precision, stream):
    return capi.lib().pbb_plan_matvec_noalloc(plan._h, C.c_void_p(w.data_ptr()), int(ncols),
                                              C.c_void_p(out.data_ptr()), int(precision),
                                              C.c_void_p(stream))


def _pair(torch, plan, ncols, precision, seed, gpu):
    """(pbb_plan_matvec output, pbb_plan_matvec_noalloc output) on the same weights."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    w = (torch.rand((ncols, plan.npoints), generator=g, dtype=torch.float64) * 2 - 1).to(gpu)
    ref = torch.zeros_like(w)
    got = torch.zeros_like(w)
    stream = torch.cuda.current_stream(gpu).cuda_stream
    plan.matvec_ptr(w.data_ptr(), ncols, ref.data_ptr(), precision, stream)
    capi.check(_noalloc(plan, w, got, ncols, precision, stream))
    torch.cuda.synchronize()
    return ref, got


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def test_noalloc_equals_plan_matvec(gpu):
    import torch
    rng = np.random.default_rng(11)
    print("noalloc: 1/r, 20000 points, level 4, order 4, 3 columns, FP64", flush=True)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 4, 4, 1e-5)
    big = pbb.Plan(ops, rng.random((20000, 3)))
    ref, got = _pair(torch, big, 3, capi.PREC_FP64, 1, gpu)
    assert _rel(got, ref) <= 1e-14, _rel(got, ref)

    print("noalloc: second call, 3000 points, level 3, 1 column, FP64", flush=True)
    ops3 = pbb.Operators(pbb.laplacian(), 1.0, 3, 4, 1e-5)
    small = pbb.Plan(ops3, rng.random((3000, 3)))
    ref, got = _pair(torch, small, 1, capi.PREC_FP64, 2, gpu)
    assert _rel(got, ref) <= 1e-14, _rel(got, ref)

    print("noalloc: f16x2 after pbb_plan_matvec split the operators", flush=True)
    ref, got = _pair(torch, big, 2, capi.PREC_MIXED_F16X2, 3, gpu)
    assert _rel(got, ref) <= 1e-14, _rel(got, ref)


def test_noalloc_rejects_what_does_not_fit(gpu):
    import torch
    rng = np.random.default_rng(12)
    print("noalloc: 20000 points x 160 columns must be refused", flush=True)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 3, 4, 1e-5)
    plan = pbb.Plan(ops, rng.random((20000, 3)))
    ncols = 160  # 3 x 20000 x 160 x 8 bytes of sorted weights alone exceed 64 MiB
    w = torch.ones((ncols, plan.npoints), dtype=torch.float64, device=gpu)
    out = torch.zeros_like(w)
    st = _noalloc(plan, w, out, ncols, capi.PREC_FP64, torch.cuda.current_stream(gpu).cuda_stream)
    assert st == capi.PBB_ERR_INVALID_ARG
    assert b"allocation-free" in capi.lib().pbb_last_error()
    torch.cuda.synchronize()
    assert float(out.abs().max()) == 0.0
