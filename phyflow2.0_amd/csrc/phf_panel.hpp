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
// work: lu_work_bytes(n) of device memory (nullptr: taken from the arena).
int lu_factor_f32(Lane *L, float *A32, int n, int32_t *ipiv, int32_t *info,
                  void *work = nullptr);
size_t lu_work_bytes(int n);
// In-place float32 Cholesky of the symmetric positive definite column-major
// n x n A32 (n <= 1024), returned in the LU form above (unit lower L D^-1,
// upper D L^T, identity pivots); dg: n floats of device memory (the
// diagonal of L).  info = column + 1 of the first pivot <= 0.  Lt
// (optional): also the strictly lower part of the unit lower factor,
// row-major with row stride ldt, zeros above it.
int chol_factor_f32(Lane *L, float *A32, int n, int32_t *ipiv, int32_t *info, float *dg,
                    float *Lt = nullptr, int ldt = 0);

// Float32 LU factors of B (rows == n), or of B^T B (rows > n: Cholesky in
// LU form, the pivoted LU if B^T B is not positive definite in float32), and the
// float32 inverse formed from them, with the float64 column-major copy of B
// for refinement residuals (arena-owned).
struct DenseFactors {
    int64_t rows = 0, n = 0;
    const double *Br = nullptr;  // B itself (row-major, caller-owned, outlives the factors)
    double *Bc = nullptr;
    float *LU = nullptr;
    float *Xinv = nullptr;       // float32 (P L U)^-1, row-major
    double *N64 = nullptr;       // B^T B (rows > n)
    void *lu_work = nullptr;
    float *dg = nullptr;         // diagonal of the Cholesky factor
    // Cholesky path: the unit lower factor row-major (Lt), its inverse
    // column- and row-major (Wc, Wr) and a product buffer (Wt), each
    // npad x npad with npad = n rounded up to 32
    float *Lt = nullptr, *Wc = nullptr, *Wr = nullptr, *Wt = nullptr;
    bool cholesky = false;       // B^T B factored by Cholesky (else pivoted LU)
    int32_t *ipiv = nullptr, *info = nullptr;
    int singular_at = 0;
};

// Factors of B; buffers already held by f for the same shape are reused.
// marks (optional, 3 events): recorded after B^T B, after the LU and after
// the inverse.  One host synchronisation (the factorisation's info).
int dense_factor_device(Lane *L, const double *B, int64_t rows, int64_t n, DenseFactors *f,
                        hipEvent_t *marks = nullptr);
// The same in two parts: enqueue everything on L's stream (no
// synchronisation; f->info on the device holds the result), then, once that
// work is complete and hinfo holds f->info, finish: the pivoted LU if the
// Cholesky met a non-positive pivot, and the error for a singular matrix.
int dense_factor_enqueue(Lane *L, const double *B, int64_t rows, int64_t n, DenseFactors *f,
                         hipEvent_t *marks = nullptr);
int dense_factor_finish(Lane *L, DenseFactors *f, int32_t hinfo);
// Refinement sweeps from x = 0, or from the x given (warm); sweeps go in
// bursts between host checks, the first of first_burst sweeps.
int dense_refine_device(Lane *L, const DenseFactors &f, const double *c, double *x,
                        SolveStats *st, bool warm = false, int first_burst = 6);
// Solve B x = c (rows == n) or min ||B x - c||_2 (rows > n) for the
// row-major float64 B on the device: float32 LU + float64 refinement.
int dense_solve_device(Lane *L, const double *B, int64_t rows, int64_t n, const double *c,
                       double *x, SolveStats *st);
// The strength-independent part of the grouped panel velocity of a fixed
// body: group centres and radii, and per panel and series term the factor
// that times the panel's strength gives its share of the group's
// multipole coefficient.  Device buffers from the arena of the lane that
// computed them (panel_series_geom_device); grp == nullptr: not computed
// (the body has too few panels for the grouped path).
struct PanelSeriesGeom {
    void *grp = nullptr;
    double2 *geo = nullptr;
    int64_t m = 0;
};
// Buffers already held by out for the same m are reused.
int panel_series_geom_device(Lane *L, int64_t m, const double *xs, const double *ys,
                             const double *xe, const double *ye, const double *cosA,
                             const double *sinA, PanelSeriesGeom *out);

// Velocity induced at (xE, yE) by the m vortex panels of strengths gamma.
// sg (optional): the body's PanelSeriesGeom, so the grouped path forms its
// series coefficients as strengths times factors instead of from the
// geometry.  order (optional): the targets in a spatially compact order,
// order[k * order_stride] the input index of the k-th (the FMM tree's
// Morton order of the same points), used instead of sorting them here.
int panel_induced_device(Lane *L, int64_t m, const double *gamma, int64_t ne, const double *xE,
                         const double *yE, const double *xs, const double *ys, const double *xe,
                         const double *ye, const double *cosA, const double *sinA, double *vx,
                         double *vy, int prec = PHF_PREC_F64,
                         const PanelSeriesGeom *sg = nullptr, const int32_t *order = nullptr,
                         int order_stride = 1);
int panel_influence_device(Lane *L, int64_t m, const double *xCP, const double *yCP,
                           const double *xs, const double *ys, const double *xe, const double *ye,
                           const double *cosA, const double *sinA, double *A);

}  // namespace phf
