// phf_remesh.hpp -- M4' remesh and float64 moment sums (device).
#pragma once

#include "phf_tree.hpp"

namespace phf {

struct RemeshOut {
    // inputs: optional caller-provided destinations (device) and capacity
    double *dst_x = nullptr, *dst_y = nullptr, *dst_w = nullptr;
    int64_t dst_cap = 0;
    bool want_indices = false;
    // outputs (device; arena-owned unless dst_* given)
    int64_t m = 0;
    double *x = nullptr, *y = nullptr, *w = nullptr;
    int64_t *i = nullptr, *j = nullptr;
    int64_t cells = 0;  // cells scanned (touched-range size)
};

// kind 0: M4' with parameter c (Regrid); kind 1: Tutty's diffusion weights
// with c = hNu = sqrt(nu dt) (diffusion_tutty).
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
