"""Python surface of PBBFMM3D's boost.python module ``FMMCompute``
(src/FMMCompute.cpp:59-74): ``Compute``, ``vecOfvec3``, ``vecOfdouble``,
``convert_to_numpy``, ``convert_to_vecOfdouble``, ``convert_to_vecOfvec3``.

``Compute(tree, target, source, weight, nCols, output)`` runs the FP64
matvec on the GPU and adds the result into ``output`` (column-major
``output[c*N + i]``), as the reference's H2_3D_Compute constructor does.
A ``precision`` argument (or the tree's ``precision`` attribute) selects a
mixed mode instead; the reference has FP64 only.
"""
from __future__ import annotations

import numpy as np

from ._capi import PREC_FP64
from .core import matvec_host

__all__ = ["Compute", "vecOfvec3", "vecOfdouble", "convert_to_numpy", "convert_to_vecOfdouble",
           "convert_to_vecOfvec3"]


class _Vec:
    """A growable FP64 buffer with list-like indexing (the role of the
    reference's std::vector wrappers)."""

    _width = 1

    def __init__(self, n: int = 0):
        self.data = np.zeros((n, self._width) if self._width > 1 else n)

    def __len__(self):
        return self.data.shape[0]

    def __getitem__(self, i):
        return self.data[i]

    def __setitem__(self, i, v):
        if isinstance(i, slice) and (i.start, i.stop, i.step) == (None, None, None):
            v = np.asarray(v, dtype=np.float64)
            if self._width > 1:
                v = v.reshape(-1, self._width)
            self.data = v.copy()
        else:
            self.data[i] = v

    def __iter__(self):
        return iter(self.data)

    def resize(self, n):
        new = np.zeros((n, self._width) if self._width > 1 else n)
        m = min(n, len(self))
        new[:m] = self.data[:m]
        self.data = new


class vecOfvec3(_Vec):
    """std::vector<vector3> (src/FMMCompute.cpp:62-64)."""
    _width = 3


class vecOfdouble(_Vec):
    """std::vector<double> (src/FMMCompute.cpp:65-67)."""
    _width = 1


def convert_to_vecOfvec3(obj, output: vecOfvec3) -> None:
    """N x 3 row-major points into a vecOfvec3 (src/FMMCompute.cpp:40-56)."""
    a = np.ascontiguousarray(np.asarray(obj, dtype=np.float64)).reshape(-1, 3)
    output.data = a.copy()


def convert_to_vecOfdouble(obj, output: vecOfdouble) -> None:
    """A flat array into a vecOfdouble (src/FMMCompute.cpp:26-38)."""
    output.data = np.asarray(obj, dtype=np.float64).reshape(-1).copy()


def convert_to_numpy(vec: vecOfdouble, obj: np.ndarray) -> None:
    """Copies a vecOfdouble into an existing float64 array
    (src/FMMCompute.cpp:11-24)."""
    flat = obj.reshape(-1)
    flat[: len(vec)] = vec.data[: flat.size]


class Compute:
    """H2_3D_Compute<myKernel>(tree, target, source, weight, nCols, output)
    (src/FMMCompute.cpp:68-70, include/compute.hpp:67-149)."""

    def __init__(self, tree, target: vecOfvec3, source: vecOfvec3, weight: vecOfdouble, nCols: int,
                 output: vecOfdouble, precision: int | None = None):
        # precision: a pbb_precision mode; None takes the tree's
        # (tree.precision, FP64 unless the caller set a mixed mode)
        if precision is None:
            precision = getattr(tree, "precision", PREC_FP64)
        tgt = target.data if isinstance(target, _Vec) else np.asarray(target, dtype=np.float64)
        src = source.data if isinstance(source, _Vec) else np.asarray(source, dtype=np.float64)
        tgt = tgt.reshape(-1, 3)
        src = src.reshape(-1, 3)
        # one point set when the sources are the targets; else separate
        # sets, weights at the sources and outputs at the targets
        same = tgt.shape == src.shape and np.array_equal(tgt, src)
        N, M = tgt.shape[0], src.shape[0]
        w = weight.data if isinstance(weight, _Vec) else np.asarray(weight, dtype=np.float64)
        w = np.asarray(w, dtype=np.float64).reshape(-1)
        if w.size < M * nCols:
            raise ValueError("weight must hold nCols * (number of sources) values")
        if output.data.size < N * nCols:
            output.resize(N * nCols)
        out = np.ascontiguousarray(output.data.reshape(-1)[: N * nCols])
        matvec_host(tree.operators(), tgt, w[: M * nCols], int(nCols), out, sources=None if same else src,
                    precision=int(precision))
        output.data.reshape(-1)[: N * nCols] = out
