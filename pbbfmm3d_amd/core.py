"""Python objects over the C ABI: kernels, operators, plans.

``Operators`` is what ``H2_3D_Tree::buildFMMTree`` produces
(H2_3D_Tree.cpp:40-66); ``Plan`` is one point set's tree, kept on the GPU
(FMMDistribute, compute.hpp:157-190); ``Plan.matvec`` is FMMCompute
(compute.hpp:192-222) on device-resident weights.  Weights and outputs are
column-major ``w[c*N + i]``, outputs accumulated, as in compute.hpp:747.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np

from . import _capi as capi


@dataclass(frozen=True)
class Kernel:
    """A kernel with a device functor (pbb_kernel_id) and the properties
    SetKernelProperty sets (kernel_Types.cpp:7-13)."""
    id: int
    symmetry: int = 1
    homogen: float = 0.0
    params: Sequence[float] = ()

    def desc(self) -> capi.KernelDesc:
        d = capi.KernelDesc()
        d.id = self.id
        d.symmetry = self.symmetry
        d.homogen = float(self.homogen)
        for i, v in enumerate(tuple(self.params)[:4]):
            d.param[i] = float(v)
        return d

    def eval(self, t, s) -> float:
        """FP64 reference formula of the functor, K(target, source)."""
        p = list(self.params) + [0.0] * (4 - len(self.params))
        d = np.asarray(s, dtype=np.float64) - np.asarray(t, dtype=np.float64)
        r2 = float(d @ d)
        i = self.id
        with np.errstate(divide="ignore"):
            if i == capi.KERNEL_LAPLACIAN: return 1.0 / np.sqrt(r2)
            if i == capi.KERNEL_ONEOVERR2: return (1.0 / np.sqrt(r2)) ** 2
            if i == capi.KERNEL_GAUSSIAN: return float(np.exp(-r2 / p[0] ** 2))
            if i == capi.KERNEL_EXPONENTIAL: return p[1] * float(np.exp(-np.sqrt(r2) / p[0]))
            if i == capi.KERNEL_SOFT_LAPLACIAN: return 1.0 / np.sqrt(r2 + p[0])
            if i == capi.KERNEL_QUADRIC: return 1.0 + r2
            if i == capi.KERNEL_INV_QUADRIC: return 1.0 / (1.0 + r2)
            if i == capi.KERNEL_LOGARITHM: return 0.0 if r2 == 0 else 0.5 * np.log(r2)
            if i == capi.KERNEL_THIN_PLATE: return 0.0 if r2 == 0 else 0.5 * r2 * np.log(r2)
            if i == capi.KERNEL_R2: return r2
            if i == capi.KERNEL_CONSTANT: return 1.0
            if i == capi.KERNEL_EXP_ANISO:
                return p[0] * float(np.exp(-np.sqrt(p[1] * d[0] ** 2 + p[2] * d[1] ** 2 + p[3] * d[2] ** 2)))
        raise ValueError(f"unknown kernel id {i}")


# The reference's built-in kernels (kernel_Types.cpp) and the parameterised
# kernels the workloads use.
def laplacian() -> Kernel:
    return Kernel(capi.KERNEL_LAPLACIAN, 1, -1.0)


def oneoverr2() -> Kernel:
    return Kernel(capi.KERNEL_ONEOVERR2, 1, -2.0)


def gaussian(length: float = 1.0) -> Kernel:
    """exp(-r^2 / length^2); length 1 is kernel_Gaussian."""
    return Kernel(capi.KERNEL_GAUSSIAN, 1, 0.0, (length,))


def exponential(length: float = 1.0, scale: float = 1.0) -> Kernel:
    """scale * exp(-r / length)."""
    return Kernel(capi.KERNEL_EXPONENTIAL, 1, 0.0, (length, scale))


def soft_laplacian(delta: float) -> Kernel:
    """1 / sqrt(r^2 + delta)."""
    return Kernel(capi.KERNEL_SOFT_LAPLACIAN, 1, 0.0, (delta,))


def quadric() -> Kernel:
    return Kernel(capi.KERNEL_QUADRIC, 1, 0.0)


def inverse_quadric() -> Kernel:
    return Kernel(capi.KERNEL_INV_QUADRIC, 1, 0.0)


def logarithm() -> Kernel:
    return Kernel(capi.KERNEL_LOGARITHM, 1, 0.0)


def thin_plate_spline() -> Kernel:
    return Kernel(capi.KERNEL_THIN_PLATE, 1, 0.0)


def r2_homogeneous() -> Kernel:
    """r^2 declared homogeneous of degree 2 (a test kernel)."""
    return Kernel(capi.KERNEL_R2, 1, 2.0)


def constant() -> Kernel:
    return Kernel(capi.KERNEL_CONSTANT, 1, 0.0)


class Operators:
    """Chebyshev weights + SVD-compressed M2L operators (host FP64),
    uploaded to the GPU when the first plan is made."""

    def __init__(self, kernel: Kernel, L: float, tree_level: int, order: int, eps: float,
                 use_chebyshev: int = 1,
                 host_eval: Optional[Callable[[np.ndarray, np.ndarray], float]] = None,
                 host_precompute: bool = False):
        """The kernel blocks and the E compression run on the GPU when one
        is visible (host_precompute=True keeps them on the host); the SVDs
        run on the host."""
        self.kernel = kernel
        self._cb = None
        L_ = capi.lib()
        h = C.c_void_p()
        if host_eval is not None:
            def _fn(ctx, t, s):
                return float(host_eval(np.ctypeslib.as_array(t, (3,)),
                                       np.ctypeslib.as_array(s, (3,))))
            self._cb = capi.HOST_KERNEL_FN(_fn)
            cb = self._cb
        else:
            cb = capi.HOST_KERNEL_FN()
        d = kernel.desc()
        capi.check(L_.pbb_ops_create_ex(C.byref(d), cb, None, float(L), int(tree_level), int(order),
                                        float(eps), int(use_chebyshev),
                                        capi.OPS_HOST_PRECOMPUTE if host_precompute else 0, C.byref(h)))
        self._h = h
        self._info()

    @classmethod
    def from_arrays(cls, kernel: Kernel, L, tree_level, order, eps, rank_f, rank_s, nsets,
                    U, VT, E) -> "Operators":
        self = cls.__new__(cls)
        self.kernel = kernel
        self._cb = None
        h = C.c_void_p()
        d = kernel.desc()
        U = np.ascontiguousarray(U, dtype=np.float64)
        VT = np.ascontiguousarray(VT, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        capi.check(capi.lib().pbb_ops_create_from_arrays(
            C.byref(d), float(L), int(tree_level), int(order), float(eps), int(rank_f),
            int(rank_s), int(nsets), U.ctypes.data_as(dp), VT.ctypes.data_as(dp),
            E.ctypes.data_as(dp), C.byref(h)))
        self._h = h
        self._info()
        return self

    @staticmethod
    def cache_paths(directory: str, kernel_type: str, order: int):
        """The reference's cache file names, {kernelType}Cheb{K,U,V}{n}.bin
        (H2_3D_Tree.cpp:81-83)."""
        return tuple(os.path.join(directory, f"{kernel_type}Cheb{m}{int(order)}.bin") for m in "KUV")

    @classmethod
    def from_cache(cls, kernel: Kernel, L, tree_level, order, eps, directory: str,
                   kernel_type: str) -> "Operators":
        """Operators read from cache files in the reference's format
        (FMMReadMatrices, H2_3D_Tree.cpp:186-239).  Raises PBBError with
        status PBB_ERR_NOT_FOUND when a file is missing."""
        self = cls.__new__(cls)
        self.kernel = kernel
        self._cb = None
        h = C.c_void_p()
        d = kernel.desc()
        k, u, v = (p.encode() for p in cls.cache_paths(directory, kernel_type, order))
        capi.check(capi.lib().pbb_ops_read_cache(C.byref(d), float(L), int(tree_level), int(order),
                                                 float(eps), k, u, v, C.byref(h)))
        self._h = h
        self._info()
        return self

    def write_cache(self, directory: str, kernel_type: str) -> None:
        """Writes the operators as the reference's cache files
        (StartPrecompute / ComputeKernelCheb, H2_3D_Tree.cpp:138-162, :424-558)."""
        k, u, v = (p.encode() for p in self.cache_paths(directory, kernel_type, self.order))
        capi.check(capi.lib().pbb_ops_write_cache(self._h, k, u, v))

    def _info(self):
        info = capi.OpsInfo()
        capi.check(capi.lib().pbb_ops_query(self._h, C.byref(info)))
        self.order = info.order
        self.n3 = info.n3
        self.rank_f = info.rank_f
        self.rank_s = info.rank_s
        self.nsets = info.nsets
        self.homogeneous = bool(info.homogeneous)
        self.tree_level = info.tree_level
        self.L = info.L
        self.homogen = info.homogen
        self.eps = info.eps

    def export(self):
        """(U, VT, E, sigma_f, sigma_s) as flat column-major arrays."""
        n3, rf, rs, ns = self.n3, self.rank_f, self.rank_s, self.nsets
        U = np.zeros(ns * n3 * rf)
        VT = np.zeros(ns * rs * n3)
        E = np.zeros(ns * 316 * rf * rs)
        sf = np.zeros(n3)
        ss = np.zeros(n3)
        dp = C.POINTER(C.c_double)
        capi.check(capi.lib().pbb_ops_export(self._h, U.ctypes.data_as(dp), VT.ctypes.data_as(dp),
                                             E.ctypes.data_as(dp), sf.ctypes.data_as(dp),
                                             ss.ctypes.data_as(dp)))
        return U, VT, E, sf, ss

    def set_mixed_tolerance(self, tol: float) -> None:
        """Re-compression tolerance of the mixed modes' operator sets
        (pbb_ops_set_mixed_tolerance; default 2e-8, 0 keeps the sets' own
        factors); before the first mixed-precision call."""
        capi.check(capi.lib().pbb_ops_set_mixed_tolerance(self._h, float(tol)))

    def set_fp64_recompressed(self, enable: bool = True) -> None:
        """Run the FP64 mode with the re-compressed operator sets too (its
        result then differs from the reference's by about the tolerance)."""
        capi.check(capi.lib().pbb_ops_set_fp64_recompressed(self._h, 1 if enable else 0))

    def mixed_ranks(self):
        """[(kf, ks, discarded relative norm)] per operator set, as the
        mixed modes run them (computed on the host)."""
        ns = self.nsets
        kf = (C.c_int32 * ns)()
        ks = (C.c_int32 * ns)()
        tail = np.zeros(ns)
        capi.check(capi.lib().pbb_ops_mixed_ranks(self._h, kf, ks, tail.ctypes.data_as(C.POINTER(C.c_double))))
        return [(kf[i], ks[i], float(tail[i])) for i in range(ns)]

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and capi._lib is not None:
            capi.lib().pbb_ops_destroy(h)
            self._h = None


def _f64_cuda(t, name: str, numel: Optional[int] = None, shape=None):
    """Checks that t is a contiguous float64 CUDA tensor (of numel values or
    of the given shape) before its pointer is handed to the kernels."""
    import torch
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda:
        raise TypeError(f"{name} must be a float64 CUDA tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} must hold {numel} values, not {t.numel()}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, not {tuple(t.shape)}")
    return t


class Plan:
    """A tree on the current GPU: of one point set (the sources are the
    targets), or of separate targets and sources.

    Calls on one plan share its device workspace: issue them on one stream
    (or order the streams), as the C ABI requires."""

    def __init__(self, ops: Operators, xyz, host_tree: bool = False, sources=None,
                 reject_outside: bool = False, expansion_route: bool = False):
        """xyz: the targets (N x 3).  sources (M x 3, optional): separate
        sources; matvec then takes weights at the sources, (ncols, M), and
        returns outputs at the targets, (ncols, N) (pbb_plan_create_ts).
        host_tree: build the tree on the host instead of the GPU (the same
        arrays; kept as the GPU build's check).  reject_outside: raise
        PBBError if a point lies outside the root cube instead of
        extrapolating.  expansion_route: keep the n^3-coefficient expansions
        of every level in the mixed modes (no proxy passes at the levels of
        rank <= 32; the same sums in another order)."""
        self.ops = ops
        dp = C.POINTER(C.c_double)
        xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        flags = ((capi.PLAN_HOST_TREE if host_tree else 0) | (capi.PLAN_REJECT_OUTSIDE if reject_outside else 0)
                 | (capi.PLAN_EXPANSION_ROUTE if expansion_route else 0))
        h = C.c_void_p()
        if sources is None:
            self.npoints = self.ntargets = self.nsources = xyz.shape[0]
            capi.check(capi.lib().pbb_plan_create_ex(ops.handle, xyz.ctypes.data_as(dp), self.npoints, flags,
                                                     C.byref(h)))
        else:
            src = np.ascontiguousarray(np.asarray(sources, dtype=np.float64).reshape(-1, 3))
            self.ntargets, self.nsources = xyz.shape[0], src.shape[0]
            self.npoints = self.ntargets + self.nsources
            capi.check(capi.lib().pbb_plan_create_ts(ops.handle, xyz.ctypes.data_as(dp), self.ntargets,
                                                     src.ctypes.data_as(dp), self.nsources, flags, C.byref(h)))
        self.separate = sources is not None
        self._h = h

    def sizes(self, transpose: bool = False):
        """(points per weight column, points per output column): (sources,
        targets) for the FMM, the reverse for the transposed FMM."""
        return (self.ntargets, self.nsources) if transpose else (self.nsources, self.ntargets)

    def info(self) -> dict:
        I = capi.PlanInfo()
        capi.check(capi.lib().pbb_plan_query(self._h, C.byref(I)))
        return {"npoints": I.npoints, "tree_level": I.tree_level,
                "boxes": [I.boxes[l] for l in range(I.tree_level + 1)],
                "m2l_pairs": I.m2l_pairs, "near_pairs": I.near_pairs,
                "max_leaf_points": I.max_leaf_points, "center": tuple(I.center),
                "outside_points": I.outside_points, "ntargets": I.ntargets, "nsources": I.nsources,
                "m2l_pairs_level": [I.m2l_pairs_level[l] for l in range(I.tree_level + 1)]}

    def update_points(self, xyz, stream: Optional[int] = None) -> None:
        """Replaces the coordinates (N x 3, caller's order) keeping the tree
        as built; raises PBBError if a point leaves its leaf box."""
        xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        if xyz.shape[0] != self.npoints:
            raise ValueError(f"xyz must hold {self.npoints} points")
        if stream is None:
            import torch
            stream = torch.cuda.current_stream().cuda_stream if torch.cuda.is_available() else 0
        capi.check(capi.lib().pbb_plan_update_points(self._h, xyz.ctypes.data_as(C.POINTER(C.c_double)),
                                                     self.npoints, C.c_void_p(stream)))

    def matvec_ptr(self, w_ptr: int, ncols: int, out_ptr: int, precision: int = capi.PREC_FP64,
                   stream: int = 0) -> None:
        """Enqueues out += A w on raw device pointers (FP64, column-major)."""
        capi.check(capi.lib().pbb_plan_matvec(self._h, C.c_void_p(w_ptr), int(ncols),
                                              C.c_void_p(out_ptr), int(precision),
                                              C.c_void_p(stream)))

    def matvec(self, w, out=None, precision: int = capi.PREC_FP64):
        """torch CUDA tensors: w (ncols, N) or (N,) float64, contiguous;
        returns out (+=) of the same shape, on torch's current stream."""
        import torch
        w, out, ncols = self._io(w, out)
        stream = torch.cuda.current_stream(w.device).cuda_stream
        self.matvec_ptr(w.data_ptr(), ncols, out.data_ptr(), precision, stream)
        return out

    def _io(self, w, out, transpose: bool = False):
        """(w, out, ncols) checked: w float64 CUDA (ncols, n_in) or (n_in,);
        out (ncols, n_out) or (n_out,) on w's device, made zero if None
        (n_in, n_out from sizes(transpose); both N for one point set)."""
        import torch
        n_in, n_out = self.sizes(transpose)
        if not isinstance(w, torch.Tensor) or w.dtype != torch.float64 or not w.is_cuda:
            raise TypeError("w must be a float64 CUDA tensor")
        w = w.contiguous()
        ncols = 1 if w.dim() == 1 else w.shape[0]
        if w.dim() > 2 or w.numel() != ncols * n_in or ncols < 1:
            raise ValueError(f"w must be (ncols, {n_in}) or ({n_in},)")
        oshape = (n_out,) if w.dim() == 1 else (ncols, n_out)
        if out is None:
            out = torch.zeros(oshape, dtype=torch.float64, device=w.device)
        else:
            _f64_cuda(out, "out", shape=oshape)
            if out.device != w.device:
                raise ValueError("out must be on w's device")
        return w, out, ncols

    def matvec_transpose_ptr(self, w_ptr: int, ncols: int, out_ptr: int,
                             precision: int = capi.PREC_FP64, stream: int = 0) -> None:
        """Enqueues out += A~^T w (the adjoint FMM) on raw device pointers."""
        capi.check(capi.lib().pbb_plan_matvec_transpose(
            self._h, C.c_void_p(w_ptr), int(ncols), C.c_void_p(out_ptr), int(precision),
            C.c_void_p(stream)))

    def matvec_transpose(self, w, out=None, precision: int = capi.PREC_FP64):
        """Like matvec, with the transposed FMM operator (separate sets:
        weights at the targets, outputs at the sources)."""
        import torch
        w, out, ncols = self._io(w, out, transpose=True)
        stream = torch.cuda.current_stream(w.device).cuda_stream
        self.matvec_transpose_ptr(w.data_ptr(), ncols, out.data_ptr(), precision, stream)
        return out

    def grad_positions(self, sigma, psi, grad=None, precision: int = capi.PREC_FP64):
        """grad (N x 3, float64 CUDA) += dL/dx for out = A~(x) sigma and
        psi = dL/dout (both (ncols, N) float64 CUDA tensors)."""
        import torch
        sigma, _, ncols = self._io(sigma, None)
        if not isinstance(psi, torch.Tensor):
            raise TypeError("psi must be a float64 CUDA tensor")
        psi = _f64_cuda(psi.contiguous(), "psi", shape=sigma.shape)
        if psi.device != sigma.device:
            raise ValueError("psi must be on sigma's device")
        if grad is None:
            grad = torch.zeros((self.npoints, 3), dtype=torch.float64, device=sigma.device)
        else:
            _f64_cuda(grad, "grad", shape=(self.npoints, 3))
            if grad.device != sigma.device:
                raise ValueError("grad must be on sigma's device")
        stream = torch.cuda.current_stream(sigma.device).cuda_stream
        capi.check(capi.lib().pbb_plan_grad_positions(
            self._h, C.c_void_p(sigma.data_ptr()), C.c_void_p(psi.data_ptr()), int(ncols),
            C.c_void_p(grad.data_ptr()), int(precision), C.c_void_p(stream)))
        return grad

    def set_workspace_limit(self, nbytes: int):
        """Caps the matvec workspace (0: device memory); larger matvecs run
        in chunks of columns."""
        capi.check(capi.lib().pbb_plan_set_workspace_limit(self._h, int(nbytes)))

    def set_timing(self, enable: bool = True):
        capi.check(capi.lib().pbb_plan_set_timing(self._h, 1 if enable else 0))

    def set_m2l_tile(self, width: int, cols_per_box: int = 32, flags: int = 0,
                     precision: int = capi.PREC_FP64):
        """M2L tile of the precision's class: width columns, cols_per_box
        weight columns per target box, flags (capi.M2L_CHUNK_MAJOR,
        capi.M2L_ROW_SPLIT, capi.M2L_WAVE_ROWS, capi.M2L_SPLIT_ACC,
        capi.M2L_DIR_GROUPS(g)); width 0 restores the built-in choice."""
        capi.check(capi.lib().pbb_plan_set_m2l_tile(self._h, int(precision), int(width), int(cols_per_box),
                                                    int(flags)))

    def set_m2l_tile_level(self, level: int, width: int, cols_per_box: int = 32, flags: int = 0,
                           precision: int = capi.PREC_FP64):
        """The M2L tile of one tree level (overrides the plan-wide tile)."""
        capi.check(capi.lib().pbb_plan_set_m2l_tile_level(self._h, int(precision), int(level), int(width),
                                                          int(cols_per_box), int(flags)))

    def m2l_tiles(self, precision: int = capi.PREC_FP64) -> dict:
        """The M2L tile in force at each level 2..L, {level: (width, cols_per_box, flags)}."""
        out = {}
        t = (C.c_int32 * 3)()
        for lvl in range(2, self.info()["tree_level"] + 1):
            capi.check(capi.lib().pbb_plan_get_m2l_tile(self._h, int(precision), lvl, t))
            out[lvl] = (t[0], t[1], t[2])
        return out

    def tune_m2l(self, weights, precision: int = capi.PREC_FP64, max_rel_diff: float = 0.0) -> dict:
        """Times the candidate M2L tiles on these weights (float64 CUDA
        tensor, (ncols, N) or (N,)), keeps the fastest and returns it as
        {"width", "cols_per_box", "flags", "rel_diff"}.  max_rel_diff > 0
        (mixed modes) keeps only candidates whose result differs from the
        FP64 matvec by at most that much (the most accurate one if none
        does); rel_diff is the kept candidate's difference."""
        import torch
        w, _, ncols = self._io(weights, None)
        best = (C.c_int32 * 3)()
        got = C.c_double(0.0)
        stream = torch.cuda.current_stream(w.device).cuda_stream
        capi.check(capi.lib().pbb_plan_tune_m2l_within(
            self._h, C.c_void_p(w.data_ptr()), int(ncols), int(precision), float(max_rel_diff),
            C.c_void_p(stream), best, C.byref(got)))
        return {"width": best[0], "cols_per_box": best[1], "flags": best[2], "rel_diff": got.value,
                "levels": self.m2l_tiles(precision)}

    def timings(self) -> dict:
        ms = np.zeros(capi.NUM_PHASES)
        capi.check(capi.lib().pbb_plan_timings(self._h, ms.ctypes.data_as(C.POINTER(C.c_double))))
        return dict(zip(capi.PHASE_NAMES, ms.tolist()))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and capi._lib is not None:
            capi.lib().pbb_plan_destroy(h)
            self._h = None


class Batch:
    """Many plans' matvecs (forward or transposed) captured into one HIP
    graph.  weights[i] / outputs[i] are float64 CUDA tensors (ncols, N_i)
    bound at capture: refill them in place between launches.  transpose is
    one flag for every entry, or a sequence with one flag per entry (a
    forward matvec and an adjoint of one plan can share a launch)."""

    def __init__(self, plans, weights, outputs, ncols: int, precision: int = capi.PREC_FP64,
                 transpose=False, overwrite: bool = True):
        if not (len(plans) == len(weights) == len(outputs)) or not plans:
            raise ValueError("plans, weights and outputs must have the same non-zero length")
        n = len(plans)
        tr = [bool(transpose)] * n if isinstance(transpose, (bool, int)) else [bool(t) for t in transpose]
        if len(tr) != n:
            raise ValueError("transpose must be one flag or one flag per entry")
        for i, (p, w, o) in enumerate(zip(plans, weights, outputs)):
            n_in, n_out = p.sizes(tr[i])
            _f64_cuda(w, f"weights[{i}]", numel=ncols * n_in)
            _f64_cuda(o, f"outputs[{i}]", numel=ncols * n_out)
        self.plans, self.weights, self.outputs = list(plans), list(weights), list(outputs)
        entries = (capi.BatchEntry * n)()
        for i, (p, w, o) in enumerate(zip(plans, weights, outputs)):
            entries[i].plan = p._h.value
            entries[i].w_dev = w.data_ptr()
            entries[i].out_dev = o.data_ptr()
            entries[i].flags = (capi.BATCH_TRANSPOSE if tr[i] else 0) | \
                (capi.BATCH_OVERWRITE if overwrite else 0)
        h = C.c_void_p()
        capi.check(capi.lib().pbb_batch_create_entries(entries, n, int(ncols), int(precision),
                                                       C.byref(h)))
        self._h = h

    @staticmethod
    def tune(plans, weights, precision: int = capi.PREC_FP64, max_rel_diff: float = 0.0) -> dict:
        """M2L tiles per (operators, columns): the autotuner runs on the
        largest plan of each operator set and its tile is set on every plan
        of that set.  Returns {operators handle: tile}."""
        groups = {}
        for p, w in zip(plans, weights):
            groups.setdefault(p.ops.handle.value, []).append((p, w))
        tiles = {}
        for key, members in groups.items():
            p0, w0 = max(members, key=lambda pw: pw[0].npoints)
            t = p0.tune_m2l(w0, precision, max_rel_diff)
            for p, _ in members:
                p.set_m2l_tile(t["width"], t["cols_per_box"], t["flags"], precision)
                for lvl, (wd, cpb, fl) in t["levels"].items():
                    if (wd, cpb, fl) != (t["width"], t["cols_per_box"], t["flags"]) and \
                            lvl <= p.info()["tree_level"]:
                        p.set_m2l_tile_level(lvl, wd, cpb, fl, precision)
            tiles[key] = t
        return tiles

    def launch(self, stream: Optional[int] = None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        capi.check(capi.lib().pbb_batch_launch(self._h, C.c_void_p(stream)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and capi._lib is not None:
            capi.lib().pbb_batch_destroy(h)
            self._h = None


def matvec_host(ops: Operators, xyz, weight, ncols: int, output=None, sources=None,
                precision: int = capi.PREC_FP64) -> np.ndarray:
    """The drop-in path (pbb_matvec, or pbb_matvec_ts with separate
    sources; pbb_matvec_prec in a mixed mode): host FP64 arrays, weights
    column-major (ncols x sources), output column-major (ncols x targets),
    accumulated."""
    if precision not in (capi.PREC_FP64, capi.PREC_MIXED, capi.PREC_MIXED_F16X2, capi.PREC_MIXED_BF16X2):
        raise ValueError(f"unknown precision mode {precision}")
    xyz = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    N = xyz.shape[0]
    src = xyz if sources is None else np.ascontiguousarray(np.asarray(sources, dtype=np.float64).reshape(-1, 3))
    M = src.shape[0]
    w = np.ascontiguousarray(np.asarray(weight, dtype=np.float64).reshape(-1))
    if w.size != M * ncols:
        raise ValueError("weight must hold ncols * nsources values")
    if output is None:
        output = np.zeros(N * ncols)
    elif not (isinstance(output, np.ndarray) and output.dtype == np.float64 and output.size == N * ncols
              and output.flags.c_contiguous):
        raise ValueError("output must be a contiguous float64 array of ncols * ntargets values")
    dp = C.POINTER(C.c_double)
    if precision != capi.PREC_FP64:
        capi.check(capi.lib().pbb_matvec_prec(ops.handle, xyz.ctypes.data_as(dp), N, src.ctypes.data_as(dp), M,
                                              w.ctypes.data_as(dp), int(ncols), output.ctypes.data_as(dp),
                                              int(precision)))
    elif sources is None:
        p = xyz.ctypes.data_as(dp)
        capi.check(capi.lib().pbb_matvec(ops.handle, p, p, N, w.ctypes.data_as(dp), int(ncols),
                                         output.ctypes.data_as(dp)))
    else:
        capi.check(capi.lib().pbb_matvec_ts(ops.handle, xyz.ctypes.data_as(dp), N, src.ctypes.data_as(dp), M,
                                            w.ctypes.data_as(dp), int(ncols), output.ctypes.data_as(dp)))
    return output
