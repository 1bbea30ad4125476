"""Drop-in for pHyFlow.cpp.panels.kernels.vortex._vortexPanel_Cython_cpu
(reference shim src/python/cpp/panels/kernels/vortex/__init__.py:24-25;
kernels src/cpp/panels/cython/vortexPanel_Cython/_vortexPanel_Cython_cpu.pyx).

    assemble_influenceMatrix(xCP, yCP, xPanelStart, yPanelStart, xPanelEnd,
                             yPanelEnd, cosAlpha, sinAlpha) -> A (M x M)     pyx:81
    inducedVelocity(gamma, xEval, yEval, xPanelStart, yPanelStart, xPanelEnd,
                    yPanelEnd, cosAlpha, sinAlpha) -> (vx, vy)               pyx:243
Both run in float64 on the GPU.
"""
import numpy as np

from ....._native import call, f64, ptr

__all__ = ["assemble_influenceMatrix", "inducedVelocity"]


def assemble_influenceMatrix(xCP, yCP, xPanelStart, yPanelStart, xPanelEnd, yPanelEnd, cosAlpha,
                             sinAlpha):
    a = [f64(v) for v in (xCP, yCP, xPanelStart, yPanelStart, xPanelEnd, yPanelEnd, cosAlpha,
                          sinAlpha)]
    M = a[0].size
    if any(v.size != M for v in a):
        raise ValueError("all panel arrays must have the same length")
    A = np.empty((M, M))
    call("phf_panel_influence", *[ptr(v) for v in a], M, ptr(A))
    return A


def inducedVelocity(gamma, xEval, yEval, xPanelStart, yPanelStart, xPanelEnd, yPanelEnd,
                    cosAlpha, sinAlpha):
    g = f64(gamma)
    xE, yE = f64(xEval), f64(yEval)
    p = [f64(v) for v in (xPanelStart, yPanelStart, xPanelEnd, yPanelEnd, cosAlpha, sinAlpha)]
    if any(v.size != g.size for v in p):
        raise ValueError("panel arrays and gamma must have the same length")
    if xE.size != yE.size:
        raise ValueError("xEval and yEval must have the same length")
    vx, vy = np.empty(xE.size), np.empty(xE.size)
    call("phf_panel_induced_velocity", ptr(g), g.size, ptr(xE), ptr(yE), xE.size,
         *[ptr(v) for v in p], ptr(vx), ptr(vy))
    return vx, vy
