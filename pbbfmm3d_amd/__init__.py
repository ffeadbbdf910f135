"""pbbfmm3d_amd -- MI355X (gfx950) black-box FMM matvec with PBBFMM3D's API.

The compute path is hand-written HIP in ``csrc/`` behind the C ABI of
``include/pbbfmm3d_amd.h``; this package binds it with ctypes.
"""
from . import _capi as capi
from .core import (Batch, Kernel, Operators, Plan, constant, exponential, gaussian, inverse_quadric,
                   laplacian, logarithm, matvec_host, oneoverr2, quadric, r2_homogeneous,
                   soft_laplacian, thin_plate_spline)
from .solve import MatrixDirect, MatrixFMM


def fmm_matvec(*args, **kwargs):
    """torch.autograd-aware matvec (see pbbfmm3d_amd.autograd)."""
    from .autograd import fmm_matvec as _f
    return _f(*args, **kwargs)


__all__ = ["capi", "fmm_matvec", "MatrixFMM", "MatrixDirect", "Batch", "Kernel", "Operators", "Plan", "matvec_host", "laplacian", "oneoverr2",
           "gaussian", "exponential", "soft_laplacian", "quadric", "inverse_quadric", "logarithm",
           "thin_plate_spline", "r2_homogeneous", "constant"]
