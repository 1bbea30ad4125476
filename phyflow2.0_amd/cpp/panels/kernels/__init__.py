"""Mirror of pHyFlow.cpp.panels.kernels (reference
src/python/cpp/panels/kernels/__init__.py), which imports its kernel
modules so that kernels.vortex._vortexPanel_Cython_cpu resolves after
``from pHyFlow.cpp.panels import kernels`` (panelSolver.py:26).  Only the
vortex panel kernels are on this path; the source-panel module
(kernels.source, panelSolver.py:134, :253) is not provided.
"""
from . import vortex

__all__ = ["vortex"]
