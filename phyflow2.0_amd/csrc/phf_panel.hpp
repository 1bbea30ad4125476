// phf_panel.hpp -- panel influence, panel velocity and the dense solve.
#pragma once

#include "phf_ops.hpp"

namespace phf {

struct SolveStats {
    int iterations = 0;
    bool converged = false;
    int singular_at = 0;
    double lu_ms = 0, refine_ms = 0;
};

// In-place float32 LU with partial pivoting of the column-major n x n A32.
int lu_factor_f32(Lane *L, float *A32, int n, int32_t *ipiv, int32_t *info);

// Solve B x = c (rows == n) or min ||B x - c||_2 (rows > n) for the
// row-major float64 B on the device: float32 LU + float64 refinement.
int dense_solve_device(Lane *L, const double *B, int64_t rows, int64_t n, const double *c,
                       double *x, SolveStats *st);

}  // namespace phf
