"""Batched evaluation of many independent blob systems (phf_batch_*).

    b = BatchSystems(sizes, sigmas, prec=PREC_F32C, order=16, leaf=64)
    for s, (x, y, w) in enumerate(systems): b.set_system(s, x, y, w)
    info = b.velocity(use_graph=True)      # one captured graph launch
    vx, vy = b.get_velocity(s)

    b.set_remesh(overlap=0.8, dt=1e-3)     # velocity + Euler step + M4' remesh
    info = b.step(use_graph=True)
    x, y, w = b.get_remeshed(s)

Every FMM and remesh phase is one launch over all systems (segment table);
each sequence is captured as a hipGraph on first use and replayed.
"""
import ctypes

import numpy as np

from ._native import (PREC_F32C, PREC_F64, BatchInfo, BatchParams, BatchRemeshParams, call,
                      f64, lib, ptr)

__all__ = ["BatchSystems", "PREC_F64", "PREC_F32C"]


class BatchSystems:
    def __init__(self, sizes, sigmas, prec=PREC_F32C, order=16, leaf=64, ws=1, k=2.0):
        sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        sigmas = f64(np.broadcast_to(np.asarray(sigmas, dtype=np.float64), sizes.shape))
        self.sizes = sizes.copy()
        bp = BatchParams(int(prec), int(order), int(leaf), int(ws), float(k))
        h = ctypes.c_void_p()
        call("phf_batch_create", ptr(sizes), ptr(sigmas), sizes.size, ctypes.byref(bp),
             ctypes.byref(h))
        self._h = h
        self.last_info = None

    def close(self):
        if getattr(self, "_h", None):
            lib().phf_batch_destroy(self._h)
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

    def set_system(self, s, x, y, w):
        x, y, w = f64(x), f64(y), f64(w)
        n = int(self.sizes[s])
        if not (x.size == y.size == w.size == n):
            raise ValueError("system %d expects %d blobs" % (s, n))
        call("phf_batch_set_system", self._h, int(s), ptr(x), ptr(y), ptr(w))

    @staticmethod
    def _info(info):
        return {"ms": info.ms, "graph": bool(info.graph), "flags": info.flags,
                "total_points": info.total_points, "levels_max": info.levels_max,
                "remeshed": info.remeshed, "spilled": info.spilled}

    def velocity(self, use_graph=True):
        info = BatchInfo()
        call("phf_batch_velocity", self._h, 1 if use_graph else 0, ctypes.byref(info))
        self.last_info = self._info(info)
        return self.last_info

    def set_remesh(self, overlap, dt=0.0, vInf=(0.0, 0.0), x0h=-1.5, y0h=-1.5, c=0.0,
                   cell_factor=2.0):
        """Lattice h_s = sigma_s*overlap with origin (x0h, y0h)*h_s
        (Blobs.redistribute: -1.5 h); forward-Euler step dt with free stream
        vInf before the remesh (dt = 0: remesh the current positions)."""
        rp = BatchRemeshParams(float(dt), float(vInf[0]), float(vInf[1]), float(overlap),
                               float(x0h), float(y0h), float(c), float(cell_factor))
        call("phf_batch_set_remesh", self._h, ctypes.byref(rp))

    def step(self, use_graph=True):
        info = BatchInfo()
        call("phf_batch_step", self._h, 1 if use_graph else 0, ctypes.byref(info))
        self.last_info = self._info(info)
        return self.last_info

    def get_remeshed(self, s):
        m = ctypes.c_int64()
        call("phf_batch_get_remeshed", self._h, int(s), ctypes.byref(m), None, None, None)
        x, y, w = np.empty(m.value), np.empty(m.value), np.empty(m.value)
        if m.value:
            call("phf_batch_get_remeshed", self._h, int(s), ctypes.byref(m), ptr(x), ptr(y),
                 ptr(w))
        return x, y, w

    def get_velocity(self, s):
        n = int(self.sizes[s])
        vx, vy = np.empty(n), np.empty(n)
        call("phf_batch_get_velocity", self._h, int(s), ptr(vx), ptr(vy))
        return vx, vy
