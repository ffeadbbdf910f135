"""The C-ABI library: loads, exports every declared entry point, and turns
bad input into status codes (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from pbbfmm3d_amd import _capi as capi

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "pbbfmm3d_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"PBB_API\s+[\w\s\*]+?\b(pbb_\w+)\s*\(", text)))


def test_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 15
    L = C.CDLL(capi.LIB_PATH)
    for nm in names:
        assert hasattr(L, nm), nm
    assert set(names) <= set(capi.exported_symbols())


def test_abi_version_and_no_device_here():
    assert capi.lib().pbb_abi_version() == 2
    assert capi.lib().pbb_device_count() >= 0


def _status(fn, *args):
    return fn(*args)


def test_invalid_arguments_return_status():
    L = capi.lib()
    h = C.c_void_p()
    d = pbb.laplacian().desc()
    none = capi.HOST_KERNEL_FN()
    assert L.pbb_ops_create(C.byref(d), none, None, 1.0, 3, 4, 1e-5, 0, C.byref(h)) == \
        capi.PBB_ERR_UNSUPPORTED  # uniform-grid branch
    assert b"Chebyshev" in L.pbb_last_error()
    assert L.pbb_ops_create(C.byref(d), none, None, 1.0, 3, 1, 1e-5, 1, C.byref(h)) == \
        capi.PBB_ERR_INVALID_ARG
    assert L.pbb_ops_create(C.byref(d), none, None, -1.0, 3, 4, 1e-5, 1, C.byref(h)) == \
        capi.PBB_ERR_INVALID_ARG
    assert L.pbb_ops_create(C.byref(d), none, None, 1.0, 0, 4, 1e-5, 1, C.byref(h)) == \
        capi.PBB_ERR_INVALID_ARG
    assert L.pbb_ops_create(None, none, None, 1.0, 3, 4, 1e-5, 1, C.byref(h)) == \
        capi.PBB_ERR_INVALID_ARG
    bad = capi.KernelDesc()
    bad.id = 99
    bad.symmetry = 1
    assert L.pbb_ops_create(C.byref(bad), none, None, 1.0, 3, 4, 1e-5, 1, C.byref(h)) == \
        capi.PBB_ERR_INVALID_ARG
    assert L.pbb_plan_matvec(None, None, 1, None, 0, None) == capi.PBB_ERR_INVALID_ARG
    assert L.pbb_plan_set_m2l_tile(None, 0, 128, 32, 0) == capi.PBB_ERR_INVALID_ARG
    assert L.pbb_plan_tune_m2l(None, None, 1, 0, None, None) == capi.PBB_ERR_INVALID_ARG
    assert L.pbb_ops_query(None, None) == capi.PBB_ERR_INVALID_ARG


def test_plan_without_device_fails_cleanly():
    if capi.lib().pbb_device_count() > 0:
        pytest.skip("a device is visible")
    ops = pbb.Operators(pbb.laplacian(), 1.0, 3, 4, 1e-5)
    with pytest.raises(capi.PBBError) as e:
        pbb.Plan(ops, np.random.default_rng(0).random((100, 3)))
    assert e.value.status == capi.PBB_ERR_NO_DEVICE


def test_ops_info():
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 5, 4, 1e-6)
    assert ops.order == 4 and ops.n3 == 64
    assert not ops.homogeneous and ops.nsets == 4  # levels 2..5
    lap = pbb.Operators(pbb.laplacian(), 1.0, 5, 4, 1e-6)
    assert lap.homogeneous and lap.nsets == 1


def test_wrappers_reject_host_tensors_before_any_call():
    # the pointer of a host or non-FP64 tensor never reaches the C ABI
    import torch
    from pbbfmm3d_amd.core import Plan, _f64_cuda
    p = Plan.__new__(Plan)
    p.npoints = p.ntargets = p.nsources = 10
    p._h = None
    w = torch.zeros(2, 10, dtype=torch.float64)
    with pytest.raises(TypeError):
        p.matvec(w)
    with pytest.raises(TypeError):
        p.matvec_transpose(w)
    with pytest.raises(TypeError):
        p.grad_positions(w, w)
    with pytest.raises(TypeError):
        p.tune_m2l(w)
    with pytest.raises(TypeError):
        _f64_cuda(w.float(), "x")
    with pytest.raises(TypeError):
        _f64_cuda(np.zeros(3), "x")


def test_matvec_prec_rejects_bad_arguments():
    # argument checks of the precision-mode drop-in run before any device work
    from pbbfmm3d_amd.core import matvec_host
    L = capi.lib()
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 3, 4, 1e-6)
    dp = C.POINTER(C.c_double)
    x = np.random.default_rng(1).random((50, 3)) - 0.5
    w = np.ones(50)
    out = np.zeros(50)
    args = (x.ctypes.data_as(dp), 50, x.ctypes.data_as(dp), 50, w.ctypes.data_as(dp), 1, out.ctypes.data_as(dp))
    assert L.pbb_matvec_prec(ops.handle, *args, 9) == capi.PBB_ERR_INVALID_ARG
    assert b"precision" in L.pbb_last_error()
    assert L.pbb_matvec_prec(ops.handle, *args, -1) == capi.PBB_ERR_INVALID_ARG
    bad = list(args)
    bad[1] = 0
    assert L.pbb_matvec_prec(ops.handle, *bad, capi.PREC_MIXED) == capi.PBB_ERR_INVALID_ARG
    bad = list(args)
    bad[5] = 0
    assert L.pbb_matvec_prec(ops.handle, *bad, capi.PREC_MIXED) == capi.PBB_ERR_INVALID_ARG
    assert L.pbb_matvec_prec(None, *args, capi.PREC_MIXED) == capi.PBB_ERR_INVALID_ARG
    with pytest.raises(ValueError):
        matvec_host(ops, x, w, 1, precision=5)
    assert not out.any()


def test_plan_flags_are_checked_before_device_work():
    # an unknown flag bit fails as an argument error; the known flags pass the
    # check (and then, without a device, fail as PBB_ERR_NO_DEVICE)
    L = capi.lib()
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 3, 4, 1e-6)
    dp = C.POINTER(C.c_double)
    x = np.random.default_rng(2).random((64, 3)) - 0.5
    h = C.c_void_p()
    assert L.pbb_plan_create_ex(ops.handle, x.ctypes.data_as(dp), 64, 8, C.byref(h)) == capi.PBB_ERR_INVALID_ARG
    assert b"flags" in L.pbb_last_error()
    if L.pbb_device_count() == 0:
        for f in (capi.PLAN_EXPANSION_ROUTE, capi.PLAN_REJECT_OUTSIDE | capi.PLAN_EXPANSION_ROUTE):
            assert L.pbb_plan_create_ex(ops.handle, x.ctypes.data_as(dp), 64, f, C.byref(h)) == capi.PBB_ERR_NO_DEVICE
