"""Resident panel bodies and the coupled vortex-blob / vortex-panel step.

Counterpart of the numerical core of pHyFlow's LagrangianSolver
(src/python/lagrangian/__lagrangianSolverClass.py:269-375, 496-677) and of
Panels (src/python/panels/__panelsClass.py): the blob/panel coupling of one
time step runs on the GPU with the blobs resident in HBM.

    body = PanelBody(xPanel, yPanel, problemType='prescribed', gTotal=0.0)
    info = step_rk2(system, vel_params(...), dt, vInf=(1, 0), panels=body,
                    remesh=dict(sigma=..., overlap=...))

The time integrator is Heun's RK2 (the reference has RK4 and Euler only).
"""
import ctypes

import numpy as np

from ._native import PanelsTiming, StepInfo, StepParams, call, f64, lib, ptr

PROBLEM = {"fixed": 0, "prescribed": 1, "moving": 2}

__all__ = ["PanelBody", "step_rk2"]


def _ends(xPanel, yPanel, xs, ys, xe, ye):
    if xPanel is not None:
        # end points of consecutive panels (reference geometry dicts,
        # examples/panels/ex01_solveNoSlip_test.py:26-38)
        xp, yp = f64(xPanel), f64(yPanel)
        xs, ys, xe, ye = xp[:-1], yp[:-1], xp[1:], yp[1:]
    return tuple(f64(v).copy() for v in (xs, ys, xe, ye))


class PanelBody:
    """Closed panel polygon(s) given by panel start/end points.

    problemType: 'prescribed' (least squares with the Kelvin row, the
    reference default), 'fixed' (A s = RHS) or 'moving' (A s = RHS with A
    rebuilt whenever the body moves, __panelsClass.py:785-798).  The given
    coordinates are the body's local coordinates; updateBody(cmGlobal,
    thetaLocal) places it (rotation about the local origin, then
    translation, __panelsClass.py:530-638) and rebuilds the influence
    matrix and its factors on the GPU."""

    def __init__(self, xPanel=None, yPanel=None, problemType="prescribed", gTotal=0.0, *,
                 xs=None, ys=None, xe=None, ye=None):
        xs, ys, xe, ye = _ends(xPanel, yPanel, xs, ys, xe, ye)
        if problemType not in PROBLEM:
            raise ValueError("problemType must be 'fixed', 'prescribed' or 'moving'")
        h = ctypes.c_void_p()
        call("phf_panels_create", ptr(xs), ptr(ys), ptr(xe), ptr(ye), xs.size,
             PROBLEM[problemType], float(gTotal), ctypes.byref(h))
        self._h = h
        self.m = xs.size
        self.problemType = problemType
        self._local = (xs, ys, xe, ye)
        self.last_update = None

    def update_geometry(self, xPanel=None, yPanel=None, *, xs=None, ys=None, xe=None, ye=None,
                        wait=True):
        """New global start/end points of the same panels: geometry,
        collocation points, influence matrix and factors are rebuilt.
        Returns (and keeps in last_update) the device times of the rebuild.
        wait=False: the rebuild is only enqueued and None is returned; the
        next use of the body (solve, step_rk2, collocation points, A, another
        update) completes it and raises its errors, and a coupled step
        evaluates the blob velocity while the factors are formed."""
        xs, ys, xe, ye = _ends(xPanel, yPanel, xs, ys, xe, ye)
        if xs.size != self.m:
            raise ValueError("the body has %d panels, got %d" % (self.m, xs.size))
        if not wait:
            call("phf_panels_update_async", self._h, ptr(xs), ptr(ys), ptr(xe), ptr(ye))
            self.last_update = None
            return None
        t = PanelsTiming()
        call("phf_panels_update", self._h, ptr(xs), ptr(ys), ptr(xe), ptr(ye), ctypes.byref(t))
        self.last_update = t.as_dict()
        return self.last_update

    def updateBody(self, cmGlobal, thetaLocal, wait=True):
        """Panels.updateBody for this body: rotate the local coordinates by
        thetaLocal (anti-clockwise) about the local origin, move them to
        cmGlobal, and rebuild (wait as in update_geometry)."""
        c, s = np.cos(thetaLocal), np.sin(thetaLocal)
        cx, cy = float(cmGlobal[0]), float(cmGlobal[1])
        xs, ys, xe, ye = self._local
        return self.update_geometry(xs=c * xs - s * ys + cx, ys=s * xs + c * ys + cy,
                                    xe=c * xe - s * ye + cx, ye=s * xe + c * ye + cy, wait=wait)

    def close(self):
        if getattr(self, "_h", None):
            lib().phf_panels_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def collocation_points(self):
        x, y = np.empty(self.m), np.empty(self.m)
        call("phf_panels_info", self._h, None, ptr(x), ptr(y), None)
        return x, y

    @property
    def A(self):
        A = np.empty((self.m, self.m))
        call("phf_panels_info", self._h, None, None, None, ptr(A))
        return A

    def solve(self, vxExternal, vyExternal):
        """Panels.solve (__panelsClass.py:715-822): strengths for the slip
        velocity (vxExternal, vyExternal) at the collocation points."""
        vx, vy = f64(vxExternal), f64(vyExternal)
        s = np.empty(self.m)
        it = ctypes.c_int32()
        call("phf_panels_solve", self._h, ptr(vx), ptr(vy), ptr(s), ctypes.byref(it))
        return s


def step_rk2(system, params, dt, vInf=(0.0, 0.0), panels=None, remesh=None, panel_update="constant"):
    """Advance a resident BlobSystem by one Heun RK2 step (+ optional M4'
    remesh; remesh = dict(sigma=, overlap=, x0=-1.5h, y0=-1.5h, c=0) for M4' with the
    Wee parameter c, or hnu=sqrt(nu dt) instead of c for Tutty's diffusion
    weights; population_control=(gThresholdLocal, gThresholdGlobal) after it).
    Returns the step info (invariants before/after the remesh, timings)."""
    sp = StepParams()
    sp.dt = float(dt)
    sp.vinf_x, sp.vinf_y = float(vInf[0]), float(vInf[1])
    sp.panel_update = 1 if panel_update == "varying" else 0
    if remesh:
        h = remesh["sigma"] * remesh.get("overlap", 1.0)
        sp.remesh = 1
        sp.remesh_sigma = float(remesh["sigma"])
        sp.remesh_overlap = float(remesh.get("overlap", 1.0))
        sp.remesh_x0 = float(remesh.get("x0", -1.5 * h))  # Blobs.redistribute bounds
        sp.remesh_y0 = float(remesh.get("y0", -1.5 * h))
        if "hnu" in remesh:  # Tutty's diffusion weights, hNu = sqrt(nu dt)
            sp.remesh_kernel = 1
            sp.remesh_c = float(remesh["hnu"])
        else:  # M4' with the Wee parameter c
            sp.remesh_c = float(remesh.get("c", 0.0))
        if remesh.get("population_control"):
            # Blobs defaults gThresholdLocal = gThresholdGlobal = 1e-8 (__blobsClass.py:105-106)
            local, glob = remesh["population_control"]
            sp.pc_enabled = 1
            sp.pc_w_threshold = float(local)
            sp.pc_total_threshold = float(glob)
    info = StepInfo()
    call("phf_system_step_rk2", system._h, ctypes.byref(params),
         panels.handle if panels is not None else None, ctypes.byref(sp), ctypes.byref(info))
    return info.as_dict()
