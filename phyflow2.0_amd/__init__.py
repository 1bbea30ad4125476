"""phyflow_amd -- MI355X (gfx950) implementation of pHyFlow's vortex-blob
velocity hot path: direct Gaussian Biot-Savart sums, the 2-D FMM with the
Oseen smoother, M4' remeshing and the vortex-panel influence build/solve.

The package directory is ``phyflow2.0_amd``; it is imported under the name
``phyflow_amd`` (see ``__graft_entry__.load_package``).  Module paths mirror
the reference package ``pHyFlow``:

    phyflow_amd.cpp.blobs.kernels.gauss   vv2parGauss_gpu, fmm2d_gpu
    phyflow_amd.cpp.blobs.remesh          m4kernel_1d
    phyflow_amd.cpp.panels.kernels.vortex assemble_influenceMatrix, inducedVelocity
    phyflow_amd.blobs.base.induced        velocity
    phyflow_amd.blobs.base.regrid         Regrid
    phyflow_amd.panels.base.panelSolver   influenceMatrix, inducedVelocity, solve

All numerics run in libphyflow_amd.so on the GPU; there is no CPU fallback:
every call raises if the library or the GPU is missing.
"""
__version__ = "0.1.0"
