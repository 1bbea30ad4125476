"""Drop-in for pHyFlow.cpp.blobs.kernels.cutoff (reference shim:
src/python/cpp/blobs/kernels/cutoff/__init__.py:24-30): the algebraic cutoff
kernel, called by velocity()/vorticity() with kernel = 0 and sigma^2
(_induced.py:585, :695, :1023, :1120).

    vv2par_gpu(xB, yB, wB, xT, yT, sigmasqr, blocksize) -> (vx, vy)
    vv2par_C_cpu(xB, yB, wB, xT, yT, sigmasqr) -> (vx, vy)
        vv2par_C_cpu.c:135-151: r2 = dx^2 + dy^2 + sigma^2, s = w/r2,
        (u, v) = inv_2pi (sum -dy s, sum dx s)
    ww2par_gpu(xB, yB, wB, xT, yT, sigmasqr, blocksize) -> w
    ww2par_C_cpu(xB, yB, wB, xT, yT, sigmasqr) -> w
        ww2par_C_cpu.c:132-144: w = inv_2pi sum w_m / sqrt(r2)

Every name runs on the GPU; ``vv2par_Cython_cpu`` is the same function.
Keyword-only extension: ``prec`` (PREC_F64 default, or PREC_F32C).
"""
import numpy as np

from ...._native import PREC_F32C, PREC_F64, call, f64, ptr

__all__ = ["vv2par_gpu", "vv2par_C_cpu", "vv2par_Cython_cpu", "ww2par_gpu", "ww2par_C_cpu",
           "PREC_F64", "PREC_F32C"]


def _inputs(xB, yB, wB, xT, yT):
    xB, yB, wB, xT, yT = (f64(a) for a in (xB, yB, wB, xT, yT))
    if not (xB.size == yB.size == wB.size):
        raise ValueError("xBlob, yBlob, wBlob must have equal lengths")
    if xT.size != yT.size:
        raise ValueError("target x and y have different lengths (%d, %d)" % (xT.size, yT.size))
    return xB, yB, wB, xT, yT


def vv2par_gpu(xB, yB, wB, xT, yT, sigmasqr, blocksize=128, *, prec=PREC_F64):
    xB, yB, wB, xT, yT = _inputs(xB, yB, wB, xT, yT)
    vx, vy = np.empty(xT.size), np.empty(xT.size)
    call("phf_vv2par_cutoff", ptr(xB), ptr(yB), ptr(wB), xB.size, ptr(xT), ptr(yT), xT.size,
         float(sigmasqr), int(prec), ptr(vx), ptr(vy))
    return vx, vy


def vv2par_C_cpu(xB, yB, wB, xT, yT, sigmasqr, *, prec=PREC_F64):
    return vv2par_gpu(xB, yB, wB, xT, yT, sigmasqr, prec=prec)


vv2par_Cython_cpu = vv2par_C_cpu


def ww2par_gpu(xB, yB, wB, xT, yT, sigmasqr, blocksize=128, *, prec=PREC_F64):
    xB, yB, wB, xT, yT = _inputs(xB, yB, wB, xT, yT)
    w = np.empty(xT.size)
    call("phf_ww2par_cutoff", ptr(xB), ptr(yB), ptr(wB), xB.size, ptr(xT), ptr(yT), xT.size,
         float(sigmasqr), int(prec), ptr(w))
    return w


def ww2par_C_cpu(xB, yB, wB, xT, yT, sigmasqr, *, prec=PREC_F64):
    return ww2par_gpu(xB, yB, wB, xT, yT, sigmasqr, prec=prec)
