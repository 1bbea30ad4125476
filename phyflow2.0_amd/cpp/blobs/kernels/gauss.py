"""Drop-in for pHyFlow.cpp.blobs.kernels.gauss (reference shim:
src/python/cpp/blobs/kernels/gauss/__init__.py:24-32).

Same names, positional signatures and return conventions as the reference
native modules; the computation runs on the GPU through libphyflow_amd.

    vv2parGauss_gpu(xB, yB, wB, xT, yT, ksigmasqr, blocksize) -> (vx, vy)
        vv2parGauss_py_gpu.cu:124-206.  blocksize is accepted and ignored
        (tiling is the library's own); sizes need no padding.
    fmm2d_gpu(xB, yB, wB, xT, yT, Ndirect, xopt, cutoff, tol) -> (qr, qi)
        fmm2d_py.cpp:59-186.  Returns the reference's (qr, qi); the caller
        maps them to velocity as (-qi/2pi, -qr/2pi) (_induced.py:334-336).
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

__all__ = ["vv2parGauss_gpu", "fmm2d_gpu", "ww2parGauss_gpu", "ww2parGauss_C_cpu",
           "PREC_F64", "PREC_F32C"]


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
