"""Mirror of pHyFlow.cpp.panels.kernels.vortex (reference shim
src/python/cpp/panels/kernels/vortex/__init__.py:24-25), which imports the
module _vortexPanel_Cython_cpu; callers reach the kernels as
kernels.vortex._vortexPanel_Cython_cpu.assemble_influenceMatrix and
.inducedVelocity (panelSolver.py:130, :249).  The two functions are also
bound at this level for code that imports them directly.
"""
from . import _vortexPanel_Cython_cpu
from ._vortexPanel_Cython_cpu import assemble_influenceMatrix, inducedVelocity

__all__ = ["_vortexPanel_Cython_cpu", "assemble_influenceMatrix", "inducedVelocity"]
