"""Drop-in for pHyFlow.blobs.base.regrid (reference:
src/python/blobs/base/regrid/_regridding.py).

    Regrid(xBlob, yBlob, wBlob, sigma, overlap, xBounds, yBounds,
           interpKernel=0, c=0.0) -> (xBlobNew, yBlobNew, wBlobNew)

Same signature, defaults and return convention as the reference (:36);
the remesh runs on the GPU (deterministic float64 sums, no atomics).  Only
the M4' kernel (interpKernel 0, blobOptions.M4PRIME_INTERPKERNEL) exists in
the reference.  ``return_indices=True`` (keyword-only, not in the
reference) also returns the integer cell indices (i, j).
"""
import ctypes

import numpy as np

from ..._native import call, f64, lib, ptr

M4PRIME_INTERPKERNEL = 0

__all__ = ["Regrid", "PopulationControl", "M4PRIME_INTERPKERNEL"]


def Regrid(xBlob, yBlob, wBlob, sigma, overlap, xBounds, yBounds, interpKernel=0, c=0.0, *,
           return_indices=False):
    if interpKernel != M4PRIME_INTERPKERNEL:
        raise ValueError("interpKernel %r not available (only M4' = 0)" % (interpKernel,))
    x, y, w = f64(xBlob), f64(yBlob), f64(wBlob)
    if not (x.size == y.size == w.size):
        raise ValueError("xBlob, yBlob, wBlob must have equal lengths")
    if x.size < 1:  # _regridding.py:110-113
        out = (x.copy(), y.copy(), w.copy())
        return out + (np.empty(0, np.int64), np.empty(0, np.int64)) if return_indices else out
    h = ctypes.c_void_p()
    call("phf_regrid_m4", ptr(x), ptr(y), ptr(w), x.size, float(sigma), float(overlap),
         float(xBounds[0]), float(yBounds[0]), float(c), ctypes.byref(h))
    try:
        m = lib().phf_blobset_size(h)
        xo, yo, wo = np.empty(m), np.empty(m), np.empty(m)
        io, jo = np.empty(m, np.int64), np.empty(m, np.int64)
        call("phf_blobset_copy", h, ptr(xo), ptr(yo), ptr(wo), ptr(io), ptr(jo))
    finally:
        lib().phf_blobset_free(h)
    return (xo, yo, wo, io, jo) if return_indices else (xo, yo, wo)


def PopulationControl(xBlob, yBlob, wBlob, wThreshold, wTotalThreshold):
    """Drop-in for PopulationControl (_regridding.py:227-330): discards blobs
    with |w| < wThreshold, dividing the threshold by 10 until the discarded
    total |w| is below wTotalThreshold; returns the kept (x, y, w) in order."""
    x, y, w = f64(xBlob), f64(yBlob), f64(wBlob)
    if not (x.size == y.size == w.size):
        raise ValueError("xBlob, yBlob, wBlob must have equal lengths")
    if x.size == 0:
        return x.copy(), y.copy(), w.copy()
    h = ctypes.c_void_p()
    call("phf_population_control", ptr(x), ptr(y), ptr(w), x.size, float(wThreshold),
         float(wTotalThreshold), ctypes.byref(h))
    try:
        m = lib().phf_blobset_size(h)
        xo, yo, wo = np.empty(m), np.empty(m), np.empty(m)
        call("phf_blobset_copy", h, ptr(xo), ptr(yo), ptr(wo), None, None)
    finally:
        lib().phf_blobset_free(h)
    return xo, yo, wo
