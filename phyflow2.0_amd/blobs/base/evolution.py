"""Diffusion by modified remeshing, drop-in for the diffusion functions of
pHyFlow.blobs.base.evolution (reference:
src/python/blobs/base/evolution/_evolution.py).

    diffusion_wee(dt, nu, h, xBlob, yBlob, wBlob, sigma, xBounds, yBounds,
                  overlap=0.5, k=2, kernel=1, interpKernel=0,
                  diffusion_method=1) -> (xBlobNew, yBlobNew, wBlobNew)
        (:138-235) Regrid with the M4' diffusion parameter c = sqrt(nu dt)/h
        (:228-232).
    diffusion_tutty(dt, nu, h, xBlob, yBlob, wBlob, sigma, xBounds, yBounds,
                    overlap=0.5, k=2, kernel=1, interpKernel=0,
                    diffusion_method=1) -> (xBlobNew, yBlobNew, wBlobNew)
        (:238-445) the same lattice, cells, order and coordinates as Regrid
        with h = sigma*overlap (:333; the ``h`` argument is not used, as in
        the reference) and Tutty's diffusion weights with hNu = sqrt(dt nu)
        (:366, :385-405): mass, centre and a variance of 2 nu dt per axis.

Both run on the GPU (deterministic float64 sums).  ``k``, ``kernel`` and
``diffusion_method`` are accepted for signature compatibility.
"""
import ctypes

import numpy as np

from ..._native import call, f64, lib, ptr
from .regrid import Regrid

__all__ = ["diffusion_wee", "diffusion_tutty"]


def diffusion_wee(dt, nu, h, xBlob, yBlob, wBlob, sigma, xBounds, yBounds, overlap=0.5, k=2,
                  kernel=1, interpKernel=0, diffusion_method=1):
    c = np.sqrt(nu * dt) / h
    return Regrid(xBlob, yBlob, wBlob, sigma, overlap, xBounds, yBounds, interpKernel, c)


def diffusion_tutty(dt, nu, h, xBlob, yBlob, wBlob, sigma, xBounds, yBounds, overlap=0.5, k=2,
                    kernel=1, interpKernel=0, diffusion_method=1, *, return_indices=False):
    x, y, w = f64(xBlob), f64(yBlob), f64(wBlob)
    if not (x.size == y.size == w.size):
        raise ValueError("xBlob, yBlob, wBlob must have equal lengths")
    if x.size < 1:  # :325-328
        out = (x.copy(), y.copy(), w.copy())
        return out + (np.empty(0, np.int64), np.empty(0, np.int64)) if return_indices else out
    hset = ctypes.c_void_p()
    call("phf_diffusion_tutty", ptr(x), ptr(y), ptr(w), x.size, float(sigma), float(overlap),
         float(xBounds[0]), float(yBounds[0]), float(dt), float(nu), ctypes.byref(hset))
    try:
        m = lib().phf_blobset_size(hset)
        xo, yo, wo = np.empty(m), np.empty(m), np.empty(m)
        io, jo = np.empty(m, np.int64), np.empty(m, np.int64)
        call("phf_blobset_copy", hset, ptr(xo), ptr(yo), ptr(wo), ptr(io), ptr(jo))
    finally:
        lib().phf_blobset_free(hset)
    return (xo, yo, wo, io, jo) if return_indices else (xo, yo, wo)
