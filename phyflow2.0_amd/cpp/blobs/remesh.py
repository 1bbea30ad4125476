"""Drop-in for pHyFlow.cpp.blobs.remesh (reference shim:
src/python/cpp/blobs/remesh/__init__.py:24-29).

    m4kernel_1d(x, c) -> w      (_interpolation_kernels_1d.cpp:21-93;
                                 returns the flattened weights, as the
                                 reference does)
"""
import numpy as np

from ..._native import call, f64, ptr

__all__ = ["m4kernel_1d"]


def m4kernel_1d(x, c=0.0):
    x = f64(np.asarray(x).ravel())
    w = np.empty(x.size)
    call("phf_m4kernel_1d", ptr(x), x.size, float(c), ptr(w))
    return w
