// phf_fmm.hpp -- FMM driver interface (device pointers, one stream).
#pragma once

#include "phf_tree.hpp"

namespace phf {

struct FmmArgs {
    const double *xb, *yb, *wb;  // sources (device, input order)
    int64_t nb;
    const double *xt, *yt;       // targets (device, input order)
    int64_t nt;
    bool same;                   // targets are the sources
    double xopt, cutoff;         // Oseen smoother (fmm2d_py.cpp:61)
    int p, ws, prec, Lhint;      // order, well-separation, precision, leaf level
    double *qr, *qi;             // device outputs in target input order
};

struct FmmInfo {
    int L = 0, p = 0, nnear = 0;
    double D = 0.0;
};

int fmm_order_from_tol(double tol);
int fmm_level(int64_t nb, int leaf, int ndirect);
int fmm_device(Lane *L, const FmmArgs &a, StageTimer &tm, bool *used, FmmInfo *info);

}  // namespace phf
