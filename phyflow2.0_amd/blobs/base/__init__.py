"""Mirror of pHyFlow.blobs.base: numerical blob operators."""
