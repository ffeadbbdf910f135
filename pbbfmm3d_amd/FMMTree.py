"""Python surface of PBBFMM3D's boost.python module ``FMMTree``
(src/FMMTree.cpp:19-39): ``vector3``, ``myKernel`` and the overridable
``Atree``, each with ``buildFMMTree()`` and ``EvaluateKernel(t, s)``.

``myKernel`` is python/mykernel_python.hpp:2-19's kernel,
0.75 exp(-sqrt(dx^2/100 + dy^2/100 + dz^2/2.25)), non-homogeneous and
symmetric.  ``Atree`` subclasses override ``EvaluateKernel`` in Python (used
by the FP64 operator precompute) and name the device functor the GPU near
field evaluates through ``device_kernel`` (a pbbfmm3d_amd.Kernel).
"""
from __future__ import annotations

import math

import numpy as np

from . import _capi as capi
from .core import Kernel, Operators

__all__ = ["vector3", "myKernel", "Atree"]


class vector3:
    """3-D point (include/bbfmm.h:29-39, src/FMMTree.cpp:99-103)."""

    def __init__(self, x: float = 0.0, y: float = 0.0, z: float = 0.0):
        self.x, self.y, self.z = float(x), float(y), float(z)

    def __iter__(self):
        return iter((self.x, self.y, self.z))


def _xyz(p):
    if isinstance(p, vector3):
        return p.x, p.y, p.z
    return float(p[0]), float(p[1]), float(p[2])


class Atree:
    """The H2_3D_Tree base (include/H2_3D_Tree.hpp:16-116) for Python
    kernels: override EvaluateKernel (and SetKernelProperty) in a subclass."""

    device_kernel: Kernel | None = None

    def __init__(self, L: float, tree_level: int, interpolation_order: int, epsilon: float,
                 use_chebyshev: int):
        self.L = float(L)
        self.tree_level = int(tree_level)
        self.interpolation_order = int(interpolation_order)
        self.epsilon = float(epsilon)
        self.use_chebyshev = int(use_chebyshev)
        self.homogen = 0.0
        self.symmetry = 1
        self.kernelType = ""
        # precision mode of FMMCompute.Compute on this tree: FP64 (the
        # reference's arithmetic) unless the caller picks a mixed mode
        self.precision = capi.PREC_FP64
        self._ops = None

    def SetKernelProperty(self):
        pass

    def EvaluateKernel(self, targetpos, sourcepos) -> float:
        raise NotImplementedError("Atree subclasses must override EvaluateKernel")

    def buildFMMTree(self):
        """ComputeWeights + PrecomputeM2L (H2_3D_Tree.cpp:40-66)."""
        self.SetKernelProperty()
        dk = self.device_kernel
        if dk is None:
            kern = Kernel(capi.KERNEL_NONE, self.symmetry, self.homogen)
        else:
            kern = Kernel(dk.id, self.symmetry, self.homogen, tuple(dk.params))
        evaluate = self.EvaluateKernel

        def host_eval(t, s):
            return evaluate(vector3(*t), vector3(*s))

        self._ops = Operators(kern, self.L, self.tree_level, self.interpolation_order, self.epsilon,
                              self.use_chebyshev, host_eval=host_eval)
        return self

    def operators(self) -> Operators:
        if self._ops is None:
            raise RuntimeError("buildFMMTree() has not been called")
        return self._ops


class myKernel(Atree):
    """0.75 exp(-sqrt(dx^2/100 + dy^2/100 + dz^2/2.25))
    (python/mykernel_python.hpp:2-19)."""

    device_kernel = Kernel(capi.KERNEL_EXP_ANISO, 1, 0.0, (0.75, 1 / 100.0, 1 / 100.0, 1 / 2.25))

    def SetKernelProperty(self):
        self.homogen = 0.0
        self.symmetry = 1
        self.kernelType = "myKernel"

    def EvaluateKernel(self, targetpos, sourcepos) -> float:
        tx, ty, tz = _xyz(targetpos)
        sx, sy, sz = _xyz(sourcepos)
        dx, dy, dz = sx - tx, sy - ty, sz - tz
        return 0.75 * math.exp(-math.sqrt(dx * dx / 100.0 + dy * dy / 100.0 + dz * dz / 2.25))

    def buildFMMTree(self):
        # the functor is the same formula: precompute without Python callbacks
        self.SetKernelProperty()
        self._ops = Operators(self.device_kernel, self.L, self.tree_level, self.interpolation_order,
                              self.epsilon, self.use_chebyshev)
        return self


def as_numpy_points(points) -> np.ndarray:
    return np.array([_xyz(p) for p in points], dtype=np.float64)
