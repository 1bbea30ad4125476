// phf_remesh.hpp -- M4' remesh and float64 moment sums (device).
#pragma once

#include "phf_tree.hpp"

namespace phf {

struct RemeshOut {
    bool want_indices = false;  // input: also return the cell indices
    // input: device bytes read back to the host with the output count
    const void *read_src = nullptr;
    size_t read_bytes = 0;
    void *read_dst = nullptr;
    // outputs (device, arena-owned)
    int64_t m = 0;
    double *x = nullptr, *y = nullptr, *w = nullptr;
    int64_t *i = nullptr, *j = nullptr;
    int64_t cells = 0;    // cells scanned (dense touched range, or the sparse list)
    bool sparse = false;  // the range was too large for the dense layout
};

// kind 0: M4' with parameter c (Regrid); kind 1: Tutty's diffusion weights
// with c = hNu = sqrt(nu dt) (diffusion_tutty).  One host synchronisation,
// at the end (it also completes copies the caller queued before).  Not
// capturable.
int remesh_device(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                  double sigma, double overlap, double x0, double y0, double c, RemeshOut *out,
                  int kind = 0);
int m4kernel_device(Lane *L, const double *x, int64_t n, double c, double *w);
// {sum w, sum w x, sum w y, sum w (x^2+y^2)} in float64 into *d_out (device)
int moments_device(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                   double4 *d_out);

// PopulationControl: compact (x, y, w) into (xo, yo, wo) (capacity n).
int population_control_device(Lane *L, const double *x, const double *y, const double *w,
                              int64_t n, double wThreshold, double wTotalThreshold, double *xo,
                              double *yo, double *wo, int64_t *m_out, double *thr_used);

}  // namespace phf
