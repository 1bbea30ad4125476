"""Device-resident blob systems (phf_system_*): the blob state of the
reference's Blobs object (src/python/blobs/__blobsClass.py:415-494) kept in
HBM between evaluations."""
import ctypes

import numpy as np

from ._native import (METHOD_DIRECT, METHOD_FMM, PREC_F32C, PREC_F64, EvalInfo, VelParams,
                      call, f64, lib, ptr)

__all__ = ["BlobSystem", "vel_params", "METHOD_FMM", "METHOD_DIRECT", "PREC_F64", "PREC_F32C"]


def vel_params(sigma, k=2.0, method=METHOD_FMM, prec=PREC_F64, order=0, leaf=0, ws=0,
               ndirect=35, tol=1.0e-6, cutoff=None, leaf_side=0.0):
    """Parameters of velocity() (_induced.py:34-36 defaults: k=2, Ndirect=35,
    tol=1e-6, cutoff=None -> 5*xopt).  leaf_side selects the FMM domain:
    0 the cost model, < 0 the bounding box, > 0 leaves at least leaf_side
    cutoffs wide (phyflow_amd.autotune chooses it with leaf, ws and order)."""
    return VelParams(int(method), int(prec), int(order), int(leaf), int(ws), int(ndirect),
                     float(sigma), float(k), float(tol), float(cutoff) if cutoff else 0.0,
                     float(leaf_side))


class BlobSystem:
    """Blobs (x, y, w) resident on the GPU."""

    def __init__(self, x, y, w):
        x, y, w = f64(x), f64(y), f64(w)
        if not (x.size == y.size == w.size):
            raise ValueError("x, y, w must have equal lengths")
        h = ctypes.c_void_p()
        call("phf_system_create", ptr(x), ptr(y), ptr(w), x.size, ctypes.byref(h))
        self._h = h
        self.last_info = None

    def close(self):
        if getattr(self, "_h", None):
            lib().phf_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def n(self):
        v = ctypes.c_int64()
        call("phf_system_size", self._h, ctypes.byref(v))
        return v.value

    def get(self):
        n = self.n
        x, y, w = np.empty(n), np.empty(n), np.empty(n)
        call("phf_system_get", self._h, ptr(x), ptr(y), ptr(w))
        return x, y, w

    def set(self, x, y, w):
        x, y, w = f64(x), f64(y), f64(w)
        if not (x.size == y.size == w.size):
            raise ValueError("x, y, w must have equal lengths")
        call("phf_system_set", self._h, ptr(x), ptr(y), ptr(w), x.size)

    def velocity(self, params, fetch=True, info=False, count_work=False, eager=False,
                 p2p_timing_only=False):
        """Velocity at the blob centres; with fetch=False the result stays
        on the device (fetch later with last_velocity()).  info=True records
        stage timings (and, with count_work, the FMM pair/link counts) in
        last_info; p2p_timing_only=True times only the near-field stage (two
        events).  The FMM launch sequence is replayed as a
        captured HIP graph while n, the parameters and the domain side are
        unchanged; eager=True launches every kernel one by one (same
        results)."""
        n = self.n
        vx = np.empty(n) if fetch else None
        vy = np.empty(n) if fetch else None
        info = info or eager or p2p_timing_only
        ei = EvalInfo() if info else None
        if ei is not None:
            ei.flags = (1 if count_work else 0) | (2 if eager else 0) | (4 if p2p_timing_only else 0)
        call("phf_system_velocity", self._h, ctypes.byref(params),
             ptr(vx) if fetch else None, ptr(vy) if fetch else None,
             ctypes.byref(ei) if info else None)
        if info:
            self.last_info = ei.as_dict()
        return (vx, vy) if fetch else None

    def last_velocity(self):
        n = self.n
        vx, vy = np.empty(n), np.empty(n)
        call("phf_system_get_velocity", self._h, ptr(vx), ptr(vy))
        return vx, vy
