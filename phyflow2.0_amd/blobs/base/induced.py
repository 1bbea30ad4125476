"""Drop-in for pHyFlow.blobs.base.induced.velocity (reference:
src/python/blobs/base/induced/_induced.py:34-166), Python 3.

    velocity(xBlob, yBlob, wBlob, sigma, k=2, kernel=1, xEval=None,
             yEval=None, hardware=0, blocksize=128, method=0, Ndirect=35,
             tol=1.0e-6, cutoff=None) -> (vx, vy)

Dispatch as in the reference (fullOption = 100*method + 10*hardware +
kernel, :138):
  method 1 (direct), kernel 1 (gauss) -> vv2parGauss with k*sigma^2 (:401, :515)
  method 0 (FMM),    kernel 1 (gauss) -> fmm2d with xopt = sqrt(alpha k sigma^2),
                                         cutoff = 5 xopt, mapped to
                                         (-qi/2pi, -qr/2pi) (:236-251, :321-336)
  method 1 (direct), kernel 0 (cutoff) -> vv2par with sigma^2 (:585, :695)
  FMM with the cutoff kernel raises, as the reference does (:158-163).
Each branch calls the same attribute path as the reference,
kernels.gauss.fmm2d_cpu / fmm2d_gpu / vv2parGauss_C_cpu / vv2parGauss_gpu
and kernels.cutoff.vv2par_C_cpu / vv2par_gpu, chosen by ``hardware``.
Differences from the reference, all deliberate:
  * every computation runs on the GPU (the *_cpu names are bound to the same
    GPU computation; the reference's CPU and GPU variants compute the same
    sums);
  * ``xEval is None`` is tested with ``is`` (the reference's ``== None`` is
    an elementwise comparison on modern NumPy and raises in ``if``) and no
    padding is needed;
Keyword-only extensions: ``prec`` (PREC_F64 default, or PREC_F32C) and the
FMM ``order``/``leaf``/``ws``.

    vorticity(xBlob, yBlob, wBlob, sigma, k=2, kernel=1, xEval=None,
              yEval=None, hardware=0, blocksize=128, method=1) -> w

restates _induced.py:702-866: direct Gaussian vorticity through
ww2parGauss with k*sigma^2 (:862, :963); the FMM method raises as in the
reference (:798-799).  The reference's GPU branch fails when xEval is None
(xEvalNew and nEval are never set, :945-966); here the blobs are the
targets in that case.
"""
import math

import numpy as np

from ...cpp.blobs import kernels
from ...cpp.blobs.kernels.gauss import PREC_F32C, PREC_F64

GAUSS_KERNEL = 1   # blobOptions.GAUSS_KERNEL
CUTOFF_KERNEL = 0  # blobOptions.CUTOFF_KERNEL
CPU_HARDWARE = 0
GPU_HARDWARE = 1
FMM_METHOD = 0
DIRECT_METHOD = 1
ALPHA = 1.2564312086261696770  # _induced.py:236

__all__ = ["velocity", "vorticity", "PREC_F64", "PREC_F32C"]


def velocity(xBlob, yBlob, wBlob, sigma, k=2, kernel=1, xEval=None, yEval=None, hardware=0,
             blocksize=128, method=0, Ndirect=35, tol=1.0e-6, cutoff=None, *, prec=PREC_F64,
             order=0, leaf=0, ws=0):
    if hardware not in (CPU_HARDWARE, GPU_HARDWARE):
        raise ValueError("hardware must be 0 (cpu) or 1 (gpu)")
    if kernel == CUTOFF_KERNEL and method == FMM_METHOD:
        raise Exception("FMM calculation of cutoff kernel not yet available")
    if kernel not in (GAUSS_KERNEL, CUTOFF_KERNEL):
        raise ValueError("unknown kernel %r" % (kernel,))
    if (xEval is None) != (yEval is None):
        raise ValueError("xEval and yEval must both be given or both be None")
    if kernel == CUTOFF_KERNEL and method == DIRECT_METHOD:  # :585, :695 (sigma^2)
        xe = xBlob if xEval is None else xEval
        ye = yBlob if yEval is None else yEval
        if hardware == CPU_HARDWARE:
            return kernels.cutoff.vv2par_C_cpu(xBlob, yBlob, wBlob, xe, ye, sigma * sigma,
                                               prec=prec)
        return kernels.cutoff.vv2par_gpu(xBlob, yBlob, wBlob, xe, ye, sigma * sigma, blocksize,
                                         prec=prec)
    if method == DIRECT_METHOD:
        xe = xBlob if xEval is None else xEval
        ye = yBlob if yEval is None else yEval
        if hardware == CPU_HARDWARE:
            return kernels.gauss.vv2parGauss_C_cpu(xBlob, yBlob, wBlob, xe, ye, k * sigma * sigma,
                                                   prec=prec)
        return kernels.gauss.vv2parGauss_gpu(xBlob, yBlob, wBlob, xe, ye, k * sigma * sigma,
                                             blocksize, prec=prec)
    if method == FMM_METHOD:
        ksigmasqr = k * sigma * sigma
        xopt = math.sqrt(ALPHA * ksigmasqr)
        if cutoff is None:
            cutoff = 5.0 * xopt
        xe = xBlob if xEval is None else xEval
        ye = yBlob if yEval is None else yEval
        fmm2d = kernels.gauss.fmm2d_cpu if hardware == CPU_HARDWARE else kernels.gauss.fmm2d_gpu
        vy_fmm, vx_fmm = fmm2d(xBlob, yBlob, wBlob, xe, ye, Ndirect, xopt, cutoff, tol, prec=prec,
                               order=order, leaf=leaf, ws=ws)
        return -vx_fmm / (2.0 * np.pi), -vy_fmm / (2.0 * np.pi)
    raise ValueError("unknown method %r" % (method,))


def vorticity(xBlob, yBlob, wBlob, sigma, k=2, kernel=1, xEval=None, yEval=None, hardware=0,
              blocksize=128, method=1, *, prec=PREC_F64):
    if hardware not in (CPU_HARDWARE, GPU_HARDWARE):
        raise ValueError("hardware must be 0 (cpu) or 1 (gpu)")
    if method == FMM_METHOD:
        raise Exception("FMM method not implemented yet")
    if method != DIRECT_METHOD:
        raise ValueError("unknown method %r" % (method,))
    if kernel not in (GAUSS_KERNEL, CUTOFF_KERNEL):
        raise ValueError("unknown kernel %r" % (kernel,))
    if (xEval is None) != (yEval is None):
        raise ValueError("xEval and yEval must both be given or both be None")
    xe = xBlob if xEval is None else xEval
    ye = yBlob if yEval is None else yEval
    if kernel == CUTOFF_KERNEL:  # :1023, :1120 (sigma^2)
        if hardware == CPU_HARDWARE:
            return kernels.cutoff.ww2par_C_cpu(xBlob, yBlob, wBlob, xe, ye, sigma * sigma,
                                               prec=prec)
        return kernels.cutoff.ww2par_gpu(xBlob, yBlob, wBlob, xe, ye, sigma * sigma, blocksize,
                                         prec=prec)
    if hardware == CPU_HARDWARE:  # :862, :963
        return kernels.gauss.ww2parGauss_C_cpu(xBlob, yBlob, wBlob, xe, ye, k * sigma * sigma,
                                               prec=prec)
    return kernels.gauss.ww2parGauss_gpu(xBlob, yBlob, wBlob, xe, ye, k * sigma * sigma,
                                         blocksize, prec=prec)
