"""Mirror of pHyFlow.cpp.panels (reference src/python/cpp/panels/__init__.py),
which imports its kernels module."""
from . import kernels

__all__ = ["kernels"]
