"""Mirror of pHyFlow.cpp.blobs.kernels: native blob kernels."""
