// phf_ops.hpp -- internal operator interfaces (device pointers, one stream).
#pragma once

#include "phf_common.hpp"
#include "phf_device.hpp"

namespace phf {

// --- pair-kernel parameters (host, float64) -------------------------------
PairParams make_gauss_params(double ksigmasqr);
PairParams make_oseen_params(double xopt, double cutoff);

// --- host <-> device helpers ------------------------------------------------
// Copy n doubles from host to a fresh arena slice (async on lane stream).
int upload(Lane *L, const double *host, int64_t n, double **dev);
int download(Lane *L, const double *dev, int64_t n, double *host);

// --- (a) all-pairs sum ----------------------------------------------------
// o0/o1 receive, for PAIR_GAUSS_EPS, the reference velocity (vx, vy);
// for PAIR_OSEEN the reference potentials (qr, qi).  All pointers device.
int direct_device(Lane *L, int kind, int prec, const PairParams &pp,
                  const double *xb, const double *yb, const double *wb,
                  int64_t nb, const double *xt, const double *yt, int64_t nt,
                  double *o0, double *o1);
// Gaussian induced vorticity (ww2parGauss) at the targets into w (device).
int vorticity_device(Lane *L, int prec, double ksigmasqr, const double *xb, const double *yb,
                     const double *wb, int64_t nb, const double *xt, const double *yt, int64_t nt,
                     double *w, bool cutoff = false);

}  // namespace phf
