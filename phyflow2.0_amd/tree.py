"""Parity view of the Morton quad-tree the FMM builds (phf_tree_build)."""
import ctypes

import numpy as np

from ._native import call, f64, ptr


def build(x, y, L):
    """Returns (keys_in_input_order, perm, start, x0, y0, inv_s) for the
    depth-L uniform quad-tree over the points (x, y)."""
    x, y = f64(x), f64(y)
    n = x.size
    keys = np.empty(n, np.uint32)
    perm = np.empty(n, np.int32)
    start = np.empty(4 ** int(L) + 1, np.int32)
    x0, y0, inv_s = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    call("phf_tree_build", ptr(x), ptr(y), n, int(L), ptr(keys), ptr(perm), ptr(start),
         ctypes.byref(x0), ctypes.byref(y0), ctypes.byref(inv_s))
    return keys, perm, start, x0.value, y0.value, inv_s.value
