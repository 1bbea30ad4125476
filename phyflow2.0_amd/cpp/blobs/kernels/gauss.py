"""Drop-in for pHyFlow.cpp.blobs.kernels.gauss (reference shim:
src/python/cpp/blobs/kernels/gauss/__init__.py:24-32).

Same names, positional signatures and return conventions as the reference
native modules; the computation runs on the GPU through libphyflow_amd.

    vv2parGauss_gpu(xB, yB, wB, xT, yT, ksigmasqr, blocksize) -> (vx, vy)
        vv2parGauss_py_gpu.cu:124-206.  blocksize is accepted and ignored
        (tiling is the library's own); sizes need no padding.
    vv2parGauss_C_cpu(xB, yB, wB, xT, yT, ksigmasqr) -> (vx, vy)
        vv2parGauss_C_cpu.c:89-160, called by velocity() with method 1,
        hardware 0 (_induced.py:401).  Same GPU computation.
    vv2parGauss_Cython_cpu(blobs, sigmasqr) -> velocities
        vv2parGauss_Cython_cpu.pyx:47-104: blobs is an (N, 3) array of
        x, y, w; the velocity at every blob centre comes back as (N, 2).
    fmm2d_gpu(xB, yB, wB, xT, yT, Ndirect, xopt, cutoff, tol) -> (qr, qi)
        fmm2d_py.cpp:59-186.  Returns the reference's (qr, qi); the caller
        maps them to velocity as (-qi/2pi, -qr/2pi) (_induced.py:334-336).
    fmm2d_cpu(xB, yB, wB, xT, yT, Ndirect, xopt, cutoff, tol) -> (qr, qi)
        fmm2d_py.cpp:189-316 (the CPU build of the same module, exported as
        fmm2d_cpu by gauss/__init__.py:27); velocity()'s default method and
        hardware reach it (_induced.py:249).  Same GPU computation as
        fmm2d_gpu: the library has only the GPU backend.
    ww2parGauss_gpu(xB, yB, wB, xT, yT, ksigmasqr, blocksize) -> w
        ww2parGauss_py_gpu.cu:104-200 (induced vorticity); the CPU name
        ww2parGauss_C_cpu(xB, yB, wB, xT, yT, ksigmasqr) (ww2parGauss_C_cpu.c:89)
        is bound to the same GPU computation.

Keyword-only extensions (not in the reference): ``prec`` (PREC_F64 default,
or PREC_F32C for float32 pair arithmetic with compensated sums), and for
the FMM ``order`` / ``leaf`` / ``ws``.
"""
import ctypes

import numpy as np

from ...._native import (PREC_F32C, PREC_F64, FmmOpts, call, f64, ptr)

__all__ = ["vv2parGauss_C_cpu", "vv2parGauss_Cython_cpu", "vv2parGauss_gpu", "fmm2d_cpu",
           "fmm2d_gpu", "ww2parGauss_C_cpu", "ww2parGauss_gpu", "PREC_F64", "PREC_F32C"]


def _check_pair(x, y, what):
    if x.size != y.size:
        raise ValueError("%s x and y have different lengths (%d, %d)" % (what, x.size, y.size))


def vv2parGauss_gpu(xB, yB, wB, xT, yT, ksigmasqr, blocksize=128, *, prec=PREC_F64):
    xB, yB, wB, xT, yT = (f64(a) for a in (xB, yB, wB, xT, yT))
    _check_pair(xB, yB, "blob")
    if wB.size != xB.size:
        raise ValueError("wBlob length %d != nBlobs %d" % (wB.size, xB.size))
    _check_pair(xT, yT, "target")
    vx = np.empty(xT.size)
    vy = np.empty(xT.size)
    call("phf_vv2par_gauss", ptr(xB), ptr(yB), ptr(wB), xB.size, ptr(xT), ptr(yT), xT.size,
         float(ksigmasqr), int(prec), ptr(vx), ptr(vy))
    return vx, vy


def vv2parGauss_C_cpu(xB, yB, wB, xT, yT, ksigmasqr, *, prec=PREC_F64):
    return vv2parGauss_gpu(xB, yB, wB, xT, yT, ksigmasqr, prec=prec)


def vv2parGauss_Cython_cpu(blobs, sigmasqr, *, prec=PREC_F64):
    b = np.asarray(blobs)
    if b.ndim != 2 or b.shape[1] != 3:
        raise ValueError("blobs must be an (N, 3) array of x, y, w; got shape %s" % (b.shape,))
    x, y, w = (f64(b[:, j]) for j in range(3))
    vx, vy = vv2parGauss_gpu(x, y, w, x, y, sigmasqr, prec=prec)
    return np.stack((vx, vy), axis=1)


def fmm2d_gpu(xB, yB, wB, xT, yT, Ndirect, xopt, cutoff, tol, *, prec=PREC_F64,
              order=0, leaf=0, ws=0):
    xB, yB, wB, xT, yT = (f64(a) for a in (xB, yB, wB, xT, yT))
    _check_pair(xB, yB, "blob")
    if wB.size != xB.size:
        raise ValueError("wBlob length %d != nBlobs %d" % (wB.size, xB.size))
    _check_pair(xT, yT, "target")
    qr = np.empty(xT.size)
    qi = np.empty(xT.size)
    opts = FmmOpts(int(order), int(leaf), int(prec), int(ws))
    call("phf_fmm2d", ptr(xB), ptr(yB), ptr(wB), xB.size, ptr(xT), ptr(yT), xT.size,
         int(Ndirect), float(xopt), float(cutoff), float(tol), ctypes.byref(opts), ptr(qr), ptr(qi))
    return qr, qi


def fmm2d_cpu(xB, yB, wB, xT, yT, Ndirect, xopt, cutoff, tol, *, prec=PREC_F64,
              order=0, leaf=0, ws=0):
    return fmm2d_gpu(xB, yB, wB, xT, yT, Ndirect, xopt, cutoff, tol, prec=prec, order=order,
                     leaf=leaf, ws=ws)


def ww2parGauss_gpu(xB, yB, wB, xT, yT, ksigmasqr, blocksize=128, *, prec=PREC_F64):
    xB, yB, wB, xT, yT = (f64(a) for a in (xB, yB, wB, xT, yT))
    _check_pair(xB, yB, "blob")
    if wB.size != xB.size:
        raise ValueError("wBlob length %d != nBlobs %d" % (wB.size, xB.size))
    _check_pair(xT, yT, "target")
    w = np.empty(xT.size)
    call("phf_ww2par_gauss", ptr(xB), ptr(yB), ptr(wB), xB.size, ptr(xT), ptr(yT), xT.size,
         float(ksigmasqr), int(prec), ptr(w))
    return w


def ww2parGauss_C_cpu(xB, yB, wB, xT, yT, ksigmasqr, *, prec=PREC_F64):
    return ww2parGauss_gpu(xB, yB, wB, xT, yT, ksigmasqr, prec=prec)
