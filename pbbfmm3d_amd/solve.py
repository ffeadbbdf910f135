"""Matrix-free operators for iterative solvers: the reference's Eigen GMRES
application (app/fmm.hpp:21-86, app/dir.hpp:20-80, app/gmres.cpp:37-64)
on a plan that is built once and reused for every product.

``MatrixFMM(ops, xyz)`` applies A~ x + shift x with the FMM approximant A~
of ops' kernel on the points xyz (the reference's MatrixFMM adds 1e4 x "for
fast convergence", fmm.hpp:83-85, and rebuilds its tree on every product;
here the tree stays on the GPU).  ``MatrixDirect`` is the same operator
with the exact FP64 direct sum (DirectCalc3D, dir.hpp:74-76).  Both expose
``matvec`` on host vectors and ``linear_operator()`` for scipy's solvers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from .core import Operators, Plan


class MatrixFMM:
    """x -> A~ x + shift * x on N points (one column per product)."""

    def __init__(self, ops: Operators, xyz, shift: float = 1e4, precision: int = capi.PREC_FP64):
        self.ops = ops
        self.plan = Plan(ops, xyz)
        self.N = self.plan.npoints
        self.shift = float(shift)
        self.precision = precision
        self.products = 0

    @property
    def shape(self):
        return (self.N, self.N)

    def matvec(self, x):
        import torch
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.N:
            raise ValueError(f"x must hold {self.N} values")
        xd = torch.tensor(x, dtype=torch.float64, device="cuda")
        y = self.plan.matvec(xd, precision=self.precision)
        if self.shift:
            y.add_(xd, alpha=self.shift)
        self.products += 1
        return y.cpu().numpy()

    def linear_operator(self):
        from scipy.sparse.linalg import LinearOperator
        return LinearOperator(self.shape, matvec=self.matvec, dtype=np.float64)


class MatrixDirect:
    """x -> K x + shift * x with the FP64 direct sum on the GPU
    (pbb_direct_sum; isinf -> 0 as DirectCalc3D, test.hpp:18-34)."""

    def __init__(self, kernel, xyz, shift: float = 1e4):
        import torch
        self.kernel = kernel
        self.xyz = torch.tensor(np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3)),
                                device="cuda")
        self.N = self.xyz.shape[0]
        self.shift = float(shift)
        self.products = 0

    @property
    def shape(self):
        return (self.N, self.N)

    def matvec(self, x):
        import torch
        x = np.asarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.N:
            raise ValueError(f"x must hold {self.N} values")
        xd = torch.tensor(x, dtype=torch.float64, device="cuda")
        y = torch.zeros_like(xd)
        d = self.kernel.desc()
        stream = torch.cuda.current_stream().cuda_stream
        capi.check(capi.lib().pbb_direct_sum(C.byref(d), C.c_void_p(self.xyz.data_ptr()), self.N,
                                             C.c_void_p(self.xyz.data_ptr()), self.N,
                                             C.c_void_p(xd.data_ptr()), 1, C.c_void_p(y.data_ptr()),
                                             C.c_void_p(stream)))
        if self.shift:
            y.add_(xd, alpha=self.shift)
        self.products += 1
        return y.cpu().numpy()

    def linear_operator(self):
        from scipy.sparse.linalg import LinearOperator
        return LinearOperator(self.shape, matvec=self.matvec, dtype=np.float64)
