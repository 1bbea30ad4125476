"""Mirror of pHyFlow.cpp.blobs (reference src/python/cpp/blobs/__init__.py),
which imports its kernels and remesh modules."""
from . import kernels, remesh

__all__ = ["kernels", "remesh"]
