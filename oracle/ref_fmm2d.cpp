// ref_fmm2d.cpp -- C entry point onto the reference's own CPU FMM library.
//
// TEST INFRASTRUCTURE ONLY (checker and CPU baseline).  Built by
// oracle/Makefile's `ref` target against the reference's unmodified sources
// where they lie (src/cpp/blobs/c++/fmm2d_py_cpu_gpu/*.cpp, compiled with
// the CPU defines of that directory's Makefile:71-77 plus -DC_CODE) into
// oracle/_ref/libref_fmm2d.so, which is never committed.
//
// The call mirrors the reference's Python binding fmm2d_py (CPU build),
// src/cpp/blobs/c++/fmm2d_py_cpu_gpu/fmm2d_py.cpp:254-313: OSEEN smoother,
// pot=1, cont=true, sortlimits {2,2,1,1,0,0}, eta {2,2}, channelheight -1,
// no panels, potentials not requested, outputs zero-filled, and tol == 0
// selecting directInteract instead of fmm2dInteract.
#include <string.h>
#include "fmm.h"
#include "direct.h"

extern "C" void ref_fmm2d(int nb, const double *xb, const double *yb,
                          const double *wb, int nt, const double *xt,
                          const double *yt, int ndirect, double xopt,
                          double cutoff, double tol, double *qr, double *qi)
{
    bool cont = true;
    double channelheight = -1;
    SMOOTHER smooth = OSEEN;
    float sortlimits[] = {2, 2, 1, 1, 0, 0};
    double eta[] = {2, 2};
    int pot = 1;
    memset(qr, 0, sizeof(double) * (size_t)nt);
    memset(qi, 0, sizeof(double) * (size_t)nt);
    if (tol == 0.0)
        directInteract(nb, xb, yb, wb, NULL, nt, xt, yt, NULL, NULL, qr, qi,
                       NULL, 0, pot, smooth, xopt, cutoff, cont, NULL, false);
    else
        fmm2dInteract(nb, xb, yb, wb, NULL, nt, xt, yt, NULL, NULL, qr, qi,
                      NULL, 0, pot, tol, ndirect, smooth, xopt, cutoff, cont,
                      NULL, false, sortlimits, eta, channelheight);
}
