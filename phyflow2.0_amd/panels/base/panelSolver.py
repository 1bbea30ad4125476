"""Drop-in for pHyFlow.panels.base.panelSolver (reference:
src/python/panels/base/panelSolver.py) plus the dense solve of
Panels.solve (src/python/panels/__panelsClass.py:715-822).

    influenceMatrix(xCP, yCP, xPanelStart, yPanelStart, xPanelEnd, yPanelEnd,
                    cosAlpha, sinAlpha, hardware=0, method=1, panelKernel=1)   :31
    inducedVelocity(strength, xPanelStart, yPanelStart, xPanelEnd, yPanelEnd,
                    cosAlpha, sinAlpha, xEval, yEval, hardware=0, method=1,
                    panelKernel=1)                                            :148
    solve(A, RHS, problemType='prescribed', L=None, gTotal=0.0)
        -> strengths, the linear algebra of Panels.solve: 'fixed' solves
        A s = RHS (scipy lu_factor/lu_solve, :961/:782); 'moving' the same
        square system, rebuilt with the body (numpy solve / gmres / bicgstab
        to the solver tolerance, :785-798: here always the direct solve);
        'prescribed' the least-squares problem [A; L] s = [RHS; -gTotal]
        (numpy lstsq, :822).
    slip_rhs(vxExt, vyExt, cosAlpha, sinAlpha) -> RHS = -(v . t)              :771

Only the constant-strength vortex panel kernel (panelKernel 1, 'csv', the
reference default) and the direct method exist here, as in the reference's
GPU-less build; ``hardware`` selects nothing: every call runs on the GPU.
"""
import ctypes

import numpy as np

from ..._native import SolveInfo, call, f64, ptr
from ...cpp.panels import kernels

__all__ = ["influenceMatrix", "inducedVelocity", "solve", "slip_rhs"]

_PANEL_KERNEL_CSV = 1
_DIRECT_METHOD = 1


def _check(method, panelKernel):
    if method != _DIRECT_METHOD:
        raise NotImplementedError("FMM :: Not Implemented!")
    if panelKernel != _PANEL_KERNEL_CSV:
        raise NotImplementedError("only the constant-strength vortex panel kernel is provided")


def influenceMatrix(xCP, yCP, xPanelStart, yPanelStart, xPanelEnd, yPanelEnd, cosAlpha, sinAlpha,
                    hardware=0, method=1, panelKernel=1):
    _check(method, panelKernel)
    return kernels.vortex._vortexPanel_Cython_cpu.assemble_influenceMatrix(  # :130
        xCP, yCP, xPanelStart, yPanelStart, xPanelEnd, yPanelEnd, cosAlpha, sinAlpha)


def inducedVelocity(strength, xPanelStart, yPanelStart, xPanelEnd, yPanelEnd, cosAlpha, sinAlpha,
                    xEval, yEval, hardware=0, method=1, panelKernel=1):
    _check(method, panelKernel)
    return kernels.vortex._vortexPanel_Cython_cpu.inducedVelocity(  # :249
        strength, xEval, yEval, xPanelStart, yPanelStart, xPanelEnd, yPanelEnd, cosAlpha, sinAlpha)


def slip_rhs(vxExt, vyExt, cosAlpha, sinAlpha):
    """RHS of the no-slip condition for vortex panels (:771)."""
    return (-f64(vxExt)) * f64(cosAlpha) + (-f64(vyExt)) * f64(sinAlpha)


def lu_factor_f32(A):
    """(LU, piv, info): float32 LU with partial pivoting of the square A on
    the GPU in the form of scipy.linalg.lu_factor (LAPACK getrf, used by
    Panels for the 'fixed' problem, __panelsClass.py:961): LU holds U and
    the unit-lower L, piv[k] the row exchanged with row k at step k."""
    A = np.asarray(A, dtype=np.float32)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("A must be square")
    n = A.shape[0]
    Af = np.asfortranarray(A)
    LU = np.empty((n, n), dtype=np.float32, order="F")
    piv = np.empty(n, dtype=np.int32)
    info = ctypes.c_int32()
    call("phf_lu_factor_f32", Af.ctypes.data, n, LU.ctypes.data, piv.ctypes.data,
         ctypes.byref(info))
    return LU, piv, info.value


def chol_factor_f32(A):
    """(LU, info): float32 Cholesky A = L L^T of the symmetric positive
    definite A on the GPU, in the getrf form of lu_factor_f32 with identity
    pivots (LU holds the unit lower L D^-1 and the upper D L^T, D = diag L);
    the factorisation of the normal matrix of the 'prescribed' least-squares
    system (the reference solves that one by lstsq, __panelsClass.py:822).
    info = k + 1 for the first pivot <= 0."""
    A = np.asarray(A, dtype=np.float32)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("A must be square")
    n = A.shape[0]
    Af = np.asfortranarray(A)
    LU = np.empty((n, n), dtype=np.float32, order="F")
    info = ctypes.c_int32()
    call("phf_chol_factor_f32", Af.ctypes.data, n, LU.ctypes.data, ctypes.byref(info))
    return LU, info.value


def dense_solve(B, c, return_info=False):
    """x solving B x = c (square) or min ||B x - c|| (tall), float32 LU +
    float64 iterative refinement on the GPU."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    c = f64(c)
    if B.ndim != 2 or B.shape[0] != c.size or B.shape[0] < B.shape[1]:
        raise ValueError("B must be rows x n with rows >= n and len(c) == rows")
    x = np.empty(B.shape[1])
    info = SolveInfo()
    call("phf_dense_solve", ptr(B), B.shape[0], B.shape[1], ptr(c), ptr(x), ctypes.byref(info))
    if return_info:
        return x, {"iterations": info.iterations, "converged": bool(info.converged),
                   "lu_ms": info.lu_ms, "refine_ms": info.refine_ms, "total_ms": info.total_ms}
    return x


def solve(A, RHS, problemType="prescribed", L=None, gTotal=0.0, return_info=False):
    A = np.ascontiguousarray(A, dtype=np.float64)
    RHS = f64(RHS)
    if problemType in ("fixed", "moving"):
        return dense_solve(A, RHS, return_info)
    if problemType == "prescribed":
        if L is None:
            raise ValueError("the 'prescribed' problem needs the panel lengths L")
        rows = np.atleast_2d(np.asarray(L, dtype=np.float64))
        g = np.atleast_1d(np.asarray(gTotal, dtype=np.float64))
        B = np.vstack([A, rows])
        c = np.concatenate([RHS, -g])
        return dense_solve(B, c, return_info)
    raise ValueError("problemType %r not provided ('fixed', 'moving' or 'prescribed')"
                     % (problemType,))
