"""Operator cache files in the reference's format (H2_3D_Tree.cpp:74-239):
written and read through the C ABI, parsed here by an independent reader
that follows the reference's fread sequence (FMMReadMatrices :186-239,
PrecomputeAvailable :102-131), and used by the C++ drop-in's buildFMMTree
the way PrecomputeM2L uses them.  Host-only: no GPU needed."""
import os
import subprocess

import numpy as np
import pytest

import pbbfmm3d_amd as pbb
from pbbfmm3d_amd._capi import PBB_ERR_INVALID_ARG, PBB_ERR_NOT_FOUND, PBBError

HERE = os.path.dirname(os.path.abspath(__file__))


def _read_reference_files(d, kernel_type, n, nsets):
    """The reference's reader, restated: K = (double L, int level, E...),
    U = (int rank_f, U...), V = (int rank_s, VT...)."""
    k, u, v = (os.path.join(d, f"{kernel_type}Cheb{m}{n}.bin") for m in "KUV")
    kb, ub, vb = (open(p, "rb").read() for p in (k, u, v))
    L = np.frombuffer(kb[:8], np.float64)[0]
    lvl = int(np.frombuffer(kb[8:12], np.int32)[0])
    rf = int(np.frombuffer(ub[:4], np.int32)[0])
    rs = int(np.frombuffer(vb[:4], np.int32)[0])
    n3 = n ** 3
    E = np.frombuffer(kb[12:], np.float64)
    U = np.frombuffer(ub[4:], np.float64)
    VT = np.frombuffer(vb[4:], np.float64)
    assert E.size == nsets * 316 * rf * rs and U.size == nsets * n3 * rf and VT.size == nsets * rs * n3
    return L, lvl, rf, rs, U, VT, E


def _export(ops):
    U, VT, E, _, _ = ops.export()
    return U.reshape(-1), VT.reshape(-1), E.reshape(-1)


def test_write_read_round_trip_per_level(tmp_path):
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 4, 4, 1e-6)
    assert ops.nsets == 3
    ops.write_cache(str(tmp_path), "gaussian")
    L, lvl, rf, rs, U, VT, E = _read_reference_files(str(tmp_path), "gaussian", 4, 3)
    assert (L, lvl, rf, rs) == (1.0, 4, ops.rank_f, ops.rank_s)
    U0, VT0, E0 = _export(ops)
    assert np.array_equal(U, U0) and np.array_equal(VT, VT0) and np.array_equal(E, E0)
    back = pbb.Operators.from_cache(pbb.gaussian(0.3), 1.0, 4, 4, 1e-6, str(tmp_path), "gaussian")
    assert (back.rank_f, back.rank_s, back.nsets) == (ops.rank_f, ops.rank_s, 3)
    for a, b in zip(_export(back), (U0, VT0, E0)):
        assert np.array_equal(a, b)


def test_stale_missing_and_truncated_files(tmp_path):
    d = str(tmp_path)
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 4, 4, 1e-6)
    ops.write_cache(d, "gaussian")
    # per-level operators belong to one root length and depth (PrecomputeAvailable)
    for L, lvl in ((1.0, 5), (2.0, 4)):
        with pytest.raises(PBBError) as e:
            pbb.Operators.from_cache(pbb.gaussian(0.3), L, lvl, 4, 1e-6, d, "gaussian")
        assert e.value.status == PBB_ERR_INVALID_ARG
    with pytest.raises(PBBError) as e:
        pbb.Operators.from_cache(pbb.gaussian(0.3), 1.0, 4, 5, 1e-6, d, "gaussian")  # other order: no files
    assert e.value.status == PBB_ERR_NOT_FOUND
    k = os.path.join(d, "gaussianChebK4.bin")
    with open(k, "r+b") as f:
        f.truncate(os.path.getsize(k) - 8)
    with pytest.raises(PBBError) as e:
        pbb.Operators.from_cache(pbb.gaussian(0.3), 1.0, 4, 4, 1e-6, d, "gaussian")
    assert e.value.status == PBB_ERR_INVALID_ARG


def test_homogeneous_kernel_one_set_any_depth(tmp_path):
    # a homogeneous kernel's operators are computed on the unit box: one set,
    # and the header's length and level are not checked
    d = str(tmp_path)
    ops = pbb.Operators(pbb.laplacian(), 1.0, 4, 4, 1e-5)
    ops.write_cache(d, "Laplacian")
    _, _, rf, rs, U, VT, E = _read_reference_files(d, "Laplacian", 4, 1)
    back = pbb.Operators.from_cache(pbb.laplacian(), 2.0, 6, 4, 1e-5, d, "Laplacian")
    assert back.nsets == 1 and (back.rank_f, back.rank_s) == (rf, rs)
    assert np.array_equal(_export(back)[2], E)


def test_cpp_dropin_build_uses_the_cache(tmp_path):
    exe = os.path.join(HERE, "cpp", "dropin_example")
    assert os.path.exists(exe), "build tests/cpp first (__graft_entry__.build())"
    if pbb.capi.lib().pbb_device_count() > 0:
        pytest.skip("a device is visible: the example would run its GPU matvecs too")
    args = [exe, "-N", "2000", "-level", "3", "-n", "4", "-cache", str(tmp_path)]
    first = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert "kernel_Laplacian: operator cache written" in first.stdout, first.stdout + first.stderr
    # without a GPU the example stops after its first kernel's precompute
    assert {f"LaplacianCheb{m}4.bin" for m in "KUV"} <= set(os.listdir(str(tmp_path)))
    second = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert "kernel_Laplacian: operator cache read" in second.stdout, second.stdout + second.stderr
    ranks = [l for l in first.stdout.splitlines() if "ranks" in l]
    assert ranks == [l for l in second.stdout.splitlines() if "ranks" in l]


def test_parameter_file_guards_the_cache(tmp_path):
    # files written here name the parameters they hold: a read with another
    # eps or kernel parameter is refused; files without the parameter file
    # (the reference's) are read as the reference reads them
    d = str(tmp_path)
    ops = pbb.Operators(pbb.gaussian(0.3), 1.0, 3, 4, 1e-6)
    ops.write_cache(d, "Gaussian")
    k = os.path.join(d, "GaussianChebK4.bin")
    assert "eps 9.9999999999999995e-07" in open(k + ".params").read()
    back = pbb.Operators.from_cache(pbb.gaussian(0.3), 1.0, 3, 4, 1e-6, d, "Gaussian")
    assert np.array_equal(_export(back)[2], _export(ops)[2])
    for kernel, eps in ((pbb.gaussian(0.3), 1e-7), (pbb.gaussian(0.5), 1e-6)):
        with pytest.raises(PBBError) as e:
            pbb.Operators.from_cache(kernel, 1.0, 3, 4, eps, d, "Gaussian")
        assert e.value.status == PBB_ERR_INVALID_ARG and "other kernel parameters" in str(e.value)
    os.remove(k + ".params")
    pbb.Operators.from_cache(pbb.gaussian(0.5), 1.0, 3, 4, 1e-7, d, "Gaussian")
