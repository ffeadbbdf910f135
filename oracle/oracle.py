"""Python side of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module; the product never imports it.  It wraps ``libpbb_oracle.so``
(pbb_oracle.c, a plain-C restatement of PBBFMM3D's compute.hpp and the
closed-form parts of H2_3D_Tree.cpp) and adds the one piece the C file does
not restate: the two SVDs of the operator precompute, done here with
numpy.linalg.svd (LAPACK) instead of the reference's dgesvd
(H2_3D_Tree.cpp:400, :471).  The reference itself cannot be built in this
environment (its headers include FFTW3's fftw3.h, which the image lacks).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpbb_oracle.so")

# kernel ids shared with include/pbbfmm3d_amd.h
LAPLACIAN, ONEOVERR2, GAUSSIAN, EXPONENTIAL, SOFT_LAPLACIAN = 1, 2, 3, 4, 5
QUADRIC, INV_QUADRIC, LOGARITHM, THIN_PLATE, R2, CONSTANT, POLY_XY = 6, 7, 8, 9, 10, 11, 12
EXP_ANISO = 13

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.pbbo_kernel_eval.restype = C.c_double
        L.pbbo_kernel_eval.argtypes = [C.c_int, _dp] + [C.c_double] * 6
        for nm in ("pbbo_cheb_nodes", "pbbo_tkz", "pbbo_kweights", "pbbo_cweights"):
            getattr(L, nm).argtypes = [C.c_int, _dp]
            getattr(L, nm).restype = None
        L.pbbo_sn.argtypes = [C.c_int, _dp, C.c_double, _dp]
        L.pbbo_sn.restype = None
        L.pbbo_ktable.argtypes = [_ip]
        L.pbbo_ktable.restype = None
        L.pbbo_kernel_blocks.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_double, _dp, _dp]
        L.pbbo_kernel_blocks.restype = None
        L.pbbo_cutoff.argtypes = [_dp, C.c_int, C.c_double]
        L.pbbo_cutoff.restype = C.c_int
        L.pbbo_compress.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.pbbo_compress.restype = None
        L.pbbo_direct.argtypes = [C.c_int, _dp, _dp, C.c_int64, _dp, C.c_int64, _dp, C.c_int, _dp]
        L.pbbo_direct.restype = None
        L.pbbo_matvec_flat.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.c_double, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp,
                                       C.c_int, _dp, C.c_int, _dp, C.c_void_p, C.c_int]
        L.pbbo_matvec_flat.restype = C.c_int
        L.pbbo_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _params(p):
    arr = np.zeros(4)
    if p is not None:
        arr[: len(p)] = p
    return arr


# ---- closed forms ----------------------------------------------------------
def cheb_nodes(n):
    out = np.zeros(n); lib().pbbo_cheb_nodes(n, _p(out)); return out


def tkz(n):
    out = np.zeros(n * n); lib().pbbo_tkz(n, _p(out)); return out


def kweights(n):
    out = np.zeros(n ** 3); lib().pbbo_kweights(n, _p(out)); return out


def cweights(n):
    out = np.zeros(2 * n * n); lib().pbbo_cweights(n, _p(out)); return out


def sn(n, x):
    T = tkz(n)
    out = np.zeros(n)
    lib().pbbo_sn(n, _p(T), float(x), _p(out))
    return out


def ktable():
    out = np.zeros(343, dtype=np.int32)
    lib().pbbo_ktable(out.ctypes.data_as(_ip))
    return out


def kernel_eval(kid, params, t, s):
    p = _params(params)
    return lib().pbbo_kernel_eval(kid, _p(p), *map(float, t), *map(float, s))


def kernel_blocks(kid, params, symm, n, box_len):
    n3 = n ** 3
    Kw = kweights(n)
    out = np.zeros(316 * n3 * n3)
    p = _params(params)
    lib().pbbo_kernel_blocks(kid, _p(p), symm, n, float(box_len), _p(Kw), _p(out))
    # out[d*n6 + i + j*n3] = K_d(i, j)  ->  array [d, j, i]
    return out.reshape(316, n3, n3)


def cutoff(sigma, eps):
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    return lib().pbbo_cutoff(_p(s), len(s), float(eps))


# ---- operators --------------------------------------------------------------
@dataclass
class Ops:
    """Operators in the layout FMMReadMatrices reads (H2_3D_Tree.cpp:186-239):
    U nsets x (n3 x rf), VT nsets x (rs x n3), E nsets x 316 x (rf x rs), all
    column-major, flattened per set."""
    kid: int
    params: np.ndarray
    symmetry: int
    homogen: float          # as declared (1/r: -1)
    L: float
    levels: int
    n: int
    eps: float
    rf: int
    rs: int
    nsets: int
    U: np.ndarray
    VT: np.ndarray
    E: np.ndarray
    sigma_f: np.ndarray = field(default_factory=lambda: np.zeros(0))
    sigma_s: np.ndarray = field(default_factory=lambda: np.zeros(0))

    @property
    def homog(self):
        return abs(self.homogen) > 0.1


def precompute(kid, params, symmetry, homogen, L, levels, n, eps) -> Ops:
    """ComputeKernelCheb (H2_3D_Tree.cpp:251-575) per operator set, with the
    sets of StartPrecompute (:138-162) and the level-2 cutoff reused."""
    n3 = n ** 3
    homog = abs(homogen) > 0.1
    nsets = 1 if homog else max(1, levels - 1)
    Us, VTs, Es = [], [], []
    rf = rs = None
    sf = ss = None
    for s in range(nsets):
        box_len = 1.0 if homog else L / 4 / 2 ** s
        blocks = kernel_blocks(kid, params, symmetry, n, box_len)  # [d, j, i]
        Kfat = blocks.transpose(2, 0, 1).reshape(n3, 316 * n3)     # [i, (d, j)]
        U0, sig_f, _ = np.linalg.svd(Kfat, full_matrices=False)
        Kthin = blocks.transpose(0, 2, 1).reshape(316 * n3, n3)    # [(d, i), j]
        _, sig_s, VT0 = np.linalg.svd(Kthin, full_matrices=False)
        if s == 0:
            rf, rs = cutoff(sig_f, eps), cutoff(sig_s, eps)
            sf, ss = sig_f, sig_s
        U = np.asfortranarray(U0[:, :rf])
        VT = np.asfortranarray(VT0[:rs, :])
        E = np.zeros(316 * rf * rs)
        Kc = np.ascontiguousarray(blocks.reshape(-1))
        Uf = U.ravel(order="F").copy()
        VTf = VT.ravel(order="F").copy()
        lib().pbbo_compress(n3, rf, rs, _p(Uf), _p(VTf), _p(Kc), _p(E))
        Us.append(Uf); VTs.append(VTf); Es.append(E)
    return Ops(kid, _params(params), symmetry, homogen, L, levels, n, eps, rf, rs, nsets,
               np.concatenate(Us), np.concatenate(VTs), np.concatenate(Es), sf, ss)


# ---- matvec / direct sum ------------------------------------------------------
def matvec(ops: Ops, xyz, w, ncols, out=None, targets=None, nbox=0):
    """H2_3D_Compute (compute.hpp:67-149): xyz (N, 3), w column-major
    (N*ncols,), returns out += A w (column-major, N*ncols).

    targets (optional indices): form only these points' outputs (the other
    entries of out are left as they are); the far field and the near field
    then run only for the boxes that hold them, which makes the oracle
    affordable at a million points.

    nbox > 0: the tree is centred on the bbox of the first nbox points only
    (the targets, compute.hpp:97-111), the others taking the same shift
    (:121-126)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    N = xyz.shape[0]
    w = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
    assert w.size == N * ncols
    if out is None:
        out = np.zeros(N * ncols)
    mask = None
    if targets is not None:
        mask = np.zeros(N, dtype=np.uint8)
        mask[np.asarray(targets, dtype=np.int64)] = 1
    U = np.ascontiguousarray(ops.U); VT = np.ascontiguousarray(ops.VT)
    E = np.ascontiguousarray(ops.E)
    p = _params(ops.params)
    rc = lib().pbbo_matvec_flat(ops.n, ops.rf, ops.rs, ops.nsets, int(ops.homog),
                                -float(ops.homogen), float(ops.L), ops.levels, ops.kid, _p(p),
                                _p(U), _p(VT), _p(E), _p(xyz), N, _p(w), ncols, _p(out),
                                None if mask is None else mask.ctypes.data_as(C.c_void_p), int(nbox))
    if rc != 0:
        raise RuntimeError(f"oracle matvec failed ({rc})")
    return out


def matvec_ts(ops: Ops, targets_xyz, sources_xyz, w, ncols):
    """f(x_i) = sum_j K(x_i, y_j) w_j for targets x (Nt, 3) and separate
    sources y (Ns, 3), w column-major (Ns*ncols,): the FMM of the points
    [targets; sources] in the tree centred on the targets' bbox (compute.hpp:
    92-127, which centres on the targets and shifts the sources alike), with
    weight zero on the targets, read at the targets.  Returns Nt*ncols
    values, column-major."""
    t = np.ascontiguousarray(targets_xyz, dtype=np.float64).reshape(-1, 3)
    s = np.ascontiguousarray(sources_xyz, dtype=np.float64).reshape(-1, 3)
    nt, ns = t.shape[0], s.shape[0]
    w = np.asarray(w, dtype=np.float64).reshape(ncols, ns)
    wu = np.zeros((ncols, nt + ns))
    wu[:, nt:] = w
    out = matvec(ops, np.vstack([t, s]), wu.reshape(-1), ncols, targets=np.arange(nt), nbox=nt)
    return np.ascontiguousarray(out.reshape(ncols, nt + ns)[:, :nt]).reshape(-1)


def direct(kid, params, tgt, src, w, ncols):
    """DirectCalc3D (test.hpp:18-34) for Nt targets: returns column-major
    (Nt*ncols,) with out[c*Nt + i] = sum_j K(t_i, s_j) w[c*Ns + j]."""
    tgt = np.ascontiguousarray(tgt, dtype=np.float64)
    src = np.ascontiguousarray(src, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64).reshape(-1)
    out = np.zeros(tgt.shape[0] * ncols)
    p = _params(params)
    lib().pbbo_direct(kid, _p(p), _p(tgt), tgt.shape[0], _p(src), src.shape[0], _p(w), ncols,
                      _p(out))
    return out


def rel_err(a, b):
    """ComputeError (test.hpp:36-47): ||a - b|| / ||b||."""
    a = np.asarray(a).ravel(); b = np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def num_threads():
    return lib().pbbo_num_threads()
