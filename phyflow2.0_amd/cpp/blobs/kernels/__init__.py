"""Mirror of pHyFlow.cpp.blobs.kernels (reference
src/python/cpp/blobs/kernels/__init__.py), which imports the cutoff and
gauss kernel modules so that ``from pHyFlow.cpp.blobs import kernels``
followed by kernels.gauss.fmm2d_gpu etc. resolves (_induced.py:32)."""
from . import cutoff, gauss

__all__ = ["cutoff", "gauss"]
