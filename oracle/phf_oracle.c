/*
 * phf_oracle.c -- CPU FP64 restatement of the pHyFlow vortex-blob hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file: only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker (or as the timed CPU
 * baseline).  The product path runs on the GPU and fails when its HIP
 * library is missing.
 *
 * Every function cites the reference lines it restates.  How each one is
 * pinned (closed forms, invariants, the compiled reference under
 * oracle/_ref, brute force) is listed per function below; the pins live in
 * tests/test_oracle.py.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).  FP
 * contraction is off so that every expression rounds exactly as the
 * reference's own scalar x86-64 code (no FMA) or numpy does.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* A.1  Direct Gaussian Biot-Savart with EPS regularisation.
 * Restates src/cpp/blobs/c/vv2parGauss_C/vv2parGauss_C_cpu.c:99-153
 * (EPS=1e-12 at :100, inv_2pi=0.5/3.141592653589793 at :102, loop :136-153).
 * Loop order, association and constants are the reference's; OpenMP over
 * targets with schedule(static) as at :136.
 * Pinned by: the single-blob closed form, linearity in circulation, zero
 * self term, pair antisymmetry, and agreement (to the EPS term) with the
 * reference's own compiled direct Oseen sum in oracle/_ref.               */
void orc_vv2par_gauss(int64_t nb, const double *xb, const double *yb,
                      const double *wb, int64_t nt, const double *xt,
                      const double *yt, double ksigmasqr, double *vx,
                      double *vy, int nthreads)
{
    const double EPS = 1e-12;
    const double inv_2pi = 0.5 / 3.141592653589793;
    int64_t k;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (k = 0; k < nt; k++) {
        double vxk = 0.0, vyk = 0.0;
        for (int64_t m = 0; m < nb; m++) {
            double dx = xt[k] - xb[m];
            double dy = yt[k] - yb[m];
            double r2 = dx * dx + dy * dy + EPS;
            double s = (wb[m] / r2) * (1.0 - exp(-r2 / ksigmasqr));
            vxk -= dy * s;
            vyk += dx * s;
        }
        vx[k] = inv_2pi * vxk;
        vy[k] = inv_2pi * vyk;
    }
}

/* ------------------------------------------------------------------------ */
/* Algebraic cutoff kernel.  Velocity restates
 * src/cpp/blobs/c/vv2par_C/vv2par_C_cpu.c:135-151 (r2 = dx*dx + dy*dy +
 * sigmasqr, s = w/r2, vx -= dy*s, vy += dx*s, scaled by inv_2pi =
 * 0.5/3.141592653589793 at :100); the companion ww2par_C_cpu.c:132-144
 * (w += w_m/sqrt(r2), scaled by inv_2pi).
 * Pinned by: single-blob closed forms u_theta = w r/(2 pi (r^2 + s2)) and
 * w/(2 pi sqrt(r^2 + s2)); linearity; the far field of the velocity tends
 * to the point vortex; swap antisymmetry of the pair velocity.            */
void orc_vv2par_cutoff(int64_t nb, const double *xb, const double *yb,
                       const double *wb, int64_t nt, const double *xt,
                       const double *yt, double sigmasqr, double *vx,
                       double *vy, int nthreads)
{
    const double inv_2pi = 0.5 / 3.141592653589793;
    int64_t k;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (k = 0; k < nt; k++) {
        double vxk = 0.0, vyk = 0.0;
        for (int64_t m = 0; m < nb; m++) {
            double dx = xt[k] - xb[m];
            double dy = yt[k] - yb[m];
            double r2 = dx * dx + dy * dy + sigmasqr;
            double s = wb[m] / r2;
            vxk -= dy * s;
            vyk += dx * s;
        }
        vx[k] = inv_2pi * vxk;
        vy[k] = inv_2pi * vyk;
    }
}

void orc_ww2par_cutoff(int64_t nb, const double *xb, const double *yb,
                       const double *wb, int64_t nt, const double *xt,
                       const double *yt, double sigmasqr, double *w, int nthreads)
{
    const double inv_2pi = 0.5 / 3.141592653589793;
    int64_t k;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (k = 0; k < nt; k++) {
        double wk = 0.0;
        for (int64_t m = 0; m < nb; m++) {
            double dx = xt[k] - xb[m];
            double dy = yt[k] - yb[m];
            wk += wb[m] / sqrt(dx * dx + dy * dy + sigmasqr);
        }
        w[k] = inv_2pi * wk;
    }
}

/* ------------------------------------------------------------------------ */
/* Induced vorticity of Gaussian blobs.
 * Restates src/cpp/blobs/c/ww2parGauss_C/ww2parGauss_C_cpu.c:89-153
 * (EPS=1e-12 at :99, inv_pi=1/3.141592653589793 at :101, loop :133-144):
 *   wk += w_m exp(-(dx*dx + dy*dy + EPS)/ks);  w_k = inv_pi*wk/ks.
 * Pinned by: the single-blob closed form, linearity, unit mass on a fine
 * lattice (sum w h^2 == total circulation), and the curl of this oracle's
 * own velocity field (orc_vv2par_gauss) by central differences.            */
void orc_ww2par_gauss(int64_t nb, const double *xb, const double *yb,
                      const double *wb, int64_t nt, const double *xt,
                      const double *yt, double ksigmasqr, double *w,
                      int nthreads)
{
    const double EPS = 1e-12;
    const double inv_pi = 1.0 / 3.141592653589793;
    int64_t k;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (k = 0; k < nt; k++) {
        double wk = 0.0;
        for (int64_t m = 0; m < nb; m++) {
            double dx = xt[k] - xb[m];
            double dy = yt[k] - yb[m];
            double r2 = dx * dx + dy * dy + EPS;
            wk += wb[m] * exp(-r2 / ksigmasqr);
        }
        w[k] = inv_pi * wk / ksigmasqr;
    }
}

/* ------------------------------------------------------------------------ */
/* A.2  Direct sum with the FMM's Oseen smoother (pot=1, cont=true).
 * Parameters restate fmm.cpp:1166-1179 (INLINECOMPLEX_DIRECT branch):
 *   c2 = cutoff^2, shape = 1.2564312086261696770/xopt^2,
 *   scale = -1/expm1(-shape*c2).
 * Pair arithmetic restates direct.cpp:1420-1485 (the SSE2DIRECT path the
 * reference build uses, OSEENSSEALTERNATIVE off-path identical):
 *   re = er-zr, ad = re^2+im^2, P = m/ad, if ad < c2: P *= -expm1(-shape*ad)*scale
 *   (P = 0 when ad == 0);  qr -= P*re;  qi += P*im.
 * Sources outer, targets inner as in direct.cpp:1418-1421, so each target's
 * sum runs over sources in ascending order.
 * Pinned against the reference's own compiled directInteract (oracle/_ref).
 */
void orc_direct_oseen(int64_t nb, const double *xb, const double *yb,
                      const double *wb, int64_t nt, const double *xt,
                      const double *yt, double xopt, double cutoff,
                      double *qr, double *qi, int nthreads)
{
    const double c2 = cutoff * cutoff;
    const double shape = 1.2564312086261696770 / (xopt * xopt);
    const double scale = -1.0 / expm1(-shape * c2);
    int64_t i;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (i = 0; i < nt; i++) {
        double sr = 0.0, si = 0.0;
        for (int64_t j = 0; j < nb; j++) {
            double re = xt[i] - xb[j];
            double im = yt[i] - yb[j];
            double ad = re * re + im * im;
            double P = wb[j] / ad;
            if (ad < c2) {
                if (ad != 0.0)
                    P *= -expm1(-shape * ad) * scale;
                else
                    P = 0.0;
            }
            sr = sr - P * re;
            si = si + P * im;
        }
        qr[i] = sr;
        qi[i] = si;
    }
}

/* ------------------------------------------------------------------------ */
/* A.4  M4' interpolation kernel, 1-D.
 * Restates src/cpp/blobs/c++/interpolation_kernels_1d/
 * _interpolation_kernels_1d.cpp:74-89 including pow() and the strict
 * inequalities (|x|==1 and |x|==2 give 0).
 * Pinned by: partition of unity sum_k M4'(u-k)=1, interpolation property
 * M4'(0)=1, M4'(+-1)=M4'(+-2)=0, moment conservation (tests).            */
static double m4p(double x, double c)
{
    double ax = fabs(x);
    if ((ax > 1.0) && (ax < 2.0))
        return 0.5 * pow(2.0 - ax, 2.0) * (1.0 - ax) -
               pow(c, 2.0) * pow(2.0 - ax, 2.0) * (1.0 - 2.0 * ax);
    else if ((ax >= 0.0) && (ax < 1.0))
        return 1.0 - 2.5 * pow(x, 2.0) + 1.5 * pow(ax, 3.0) -
               pow(c, 2.0) * (2.0 - 9.0 * pow(x, 2) + 6.0 * pow(ax, 3.0));
    return 0.0;
}

void orc_m4kernel_1d(int64_t n, const double *x, double c, double *w)
{
    for (int64_t k = 0; k < n; k++) w[k] = m4p(x[k], c);
}

/* ------------------------------------------------------------------------ */
/* A.4  M4' remesh (Regrid).
 * Restates src/python/blobs/base/regrid/_regridding.py:118-218:
 *   iL = int64(ceil(((x - x0) - 2.0*h + 0.5*h)/h)) - 1        (:157-158)
 *   slot s: i = iL + s%4, j = jL + s/4                          (:170-172)
 *   u = ((i + 0.5) + x0/h) - x/h                                (:183)
 *   value = (M4'(ux)*M4'(uy))*w                                 (:183-184)
 *   duplicates summed, explicit zeros kept, output row-major (i, then j)
 *   via csr_matrix(...).tocoo()                                 (:200-211)
 *   x_new = (i*h + 0.5*h) + x0                                  (:217-218)
 * Within one cell the contributions are summed in input order (slot-major,
 * entry e = s*n + b), which is the order of a stable sort of the COO
 * entries; scipy's csr_sort_indices is unstable, so the reference's own
 * order inside a cell is unspecified and agrees only to rounding.
 * Pinned by the conservation laws of M4' (sum w, sum w x, sum w y,
 * sum w (x^2+y^2) for c=0), on-lattice blobs mapping to themselves, and
 * brute-force cell sets on tiny inputs (tests).                         */
typedef struct {
    int64_t i, j;
    int64_t e;
    double v;
} rm_entry;

static int rm_cmp(const void *a, const void *b)
{
    const rm_entry *p = (const rm_entry *)a, *q = (const rm_entry *)b;
    if (p->i != q->i) return p->i < q->i ? -1 : 1;
    if (p->j != q->j) return p->j < q->j ? -1 : 1;
    if (p->e != q->e) return p->e < q->e ? -1 : 1;
    return 0;
}

/* Returns the number of output cells, or -1 on allocation failure.
 * Outputs are malloc'ed and handed to the caller (free with orc_free). */
/* Tutty's diffusion weights, src/python/blobs/base/evolution/_evolution.py:
 * 385-405: for delta = (x - x_i)/h with x_i the node right of the base cell,
 * a blend (1 - delta, delta) of two 3-point stencils centred on nodes i and
 * i+1, each with mass 1, the particle's centre and variance 2 (hNu/h)^2.
 * F[a] is the weight of node i - 1 + a.                                    */
static void tutty_weights(double delta, double c2, double F[4])
{
    const double d1 = delta - 1.0;
    const double f = (1.0 - 2.0 * c2) - delta * delta;
    const double g = (1.0 - 2.0 * c2) - d1 * d1;
    F[0] = (1.0 - delta) * (0.5 * ((1.0 - f) - delta));
    F[1] = (1.0 - delta) * f + delta * (0.5 * ((1.0 - g) - d1));
    F[2] = (1.0 - delta) * (0.5 * ((1.0 - f) + delta)) + delta * g;
    F[3] = delta * (0.5 * ((1.0 - g) + d1));
}

/* kind 0: M4' with parameter c (Regrid); kind 1: Tutty with hnu = sqrt(nu dt)
 * (diffusion_tutty, _evolution.py:325-445: same base cell, stencil, COO
 * summation and output coordinates as Regrid, different weights).         */
static int64_t regrid_core(int kind, int64_t n, const double *x, const double *y,
                           const double *w, double sigma, double overlap,
                           double x0, double y0, double c, double hnu, double **xo,
                           double **yo, double **wo, int64_t **io, int64_t **jo)
{
    *xo = *yo = *wo = NULL;
    *io = *jo = NULL;
    if (n <= 0) return 0;
    const double h = sigma * overlap;
    const double h2 = 2.0 * h, hh = 0.5 * h;
    const double x0h = x0 / h, y0h = y0 / h;
    const double r = hnu / h, c2 = r * r;
    rm_entry *E = (rm_entry *)malloc(sizeof(rm_entry) * 16 * (size_t)n);
    if (!E) return -1;
    for (int64_t b = 0; b < n; b++) {
        int64_t iL = (int64_t)ceil((((x[b] - x0) - h2) + hh) / h) - 1;
        int64_t jL = (int64_t)ceil((((y[b] - y0) - h2) + hh) / h) - 1;
        double xbh = x[b] / h, ybh = y[b] / h;
        double F[4], G[4];
        if (kind == 1) {
            const double xc = ((double)(iL + 1) * h + hh) + x0;
            const double yc = ((double)(jL + 1) * h + hh) + y0;
            tutty_weights((x[b] - xc) / h, c2, F);
            tutty_weights((y[b] - yc) / h, c2, G);
        }
        for (int s = 0; s < 16; s++) {
            int64_t i = iL + (s % 4), j = jL + (s / 4);
            rm_entry *en = &E[(size_t)s * n + b];
            en->i = i;
            en->j = j;
            en->e = (int64_t)s * n + b;
            if (kind == 1) {
                en->v = (F[s % 4] * G[s / 4]) * w[b];
            } else {
                double ux = (((double)i + 0.5) + x0h) - xbh;
                double uy = (((double)j + 0.5) + y0h) - ybh;
                en->v = (m4p(ux, c) * m4p(uy, c)) * w[b];
            }
        }
    }
    size_t ne = 16 * (size_t)n;
    qsort(E, ne, sizeof(rm_entry), rm_cmp);
    int64_t m = 0;
    for (size_t k = 0; k < ne; k++)
        if (k == 0 || E[k].i != E[k - 1].i || E[k].j != E[k - 1].j) m++;
    double *X = (double *)malloc(sizeof(double) * m);
    double *Y = (double *)malloc(sizeof(double) * m);
    double *W = (double *)malloc(sizeof(double) * m);
    int64_t *I = (int64_t *)malloc(sizeof(int64_t) * m);
    int64_t *J = (int64_t *)malloc(sizeof(int64_t) * m);
    if (!X || !Y || !W || !I || !J) {
        free(E); free(X); free(Y); free(W); free(I); free(J);
        return -1;
    }
    int64_t o = -1;
    for (size_t k = 0; k < ne; k++) {
        if (k == 0 || E[k].i != E[k - 1].i || E[k].j != E[k - 1].j) {
            o++;
            I[o] = E[k].i;
            J[o] = E[k].j;
            W[o] = E[k].v;
        } else {
            W[o] = W[o] + E[k].v;
        }
    }
    for (int64_t k = 0; k < m; k++) {
        X[k] = ((double)I[k] * h + hh) + x0;
        Y[k] = ((double)J[k] * h + hh) + y0;
    }
    free(E);
    *xo = X; *yo = Y; *wo = W; *io = I; *jo = J;
    return m;
}

int64_t orc_regrid_m4(int64_t n, const double *x, const double *y,
                      const double *w, double sigma, double overlap,
                      double x0, double y0, double c, double **xo,
                      double **yo, double **wo, int64_t **io, int64_t **jo)
{
    return regrid_core(0, n, x, y, w, sigma, overlap, x0, y0, c, 0.0, xo, yo, wo, io, jo);
}

/* diffusion_tutty (_evolution.py:238-445) with h = sigma*overlap (:333) and
 * hNu = sqrt(dt*nu) (:366).  Pinned by: conservation of sum w, sum w x,
 * sum w y and growth of sum w (x^2 + y^2) by exactly 4 nu dt sum w (each
 * axis gains variance 2 nu dt); hnu = 0 on a lattice point maps a blob to
 * itself; the same cell set and order as Regrid.                          */
int64_t orc_regrid_tutty(int64_t n, const double *x, const double *y,
                         const double *w, double sigma, double overlap,
                         double x0, double y0, double hnu, double **xo,
                         double **yo, double **wo, int64_t **io, int64_t **jo)
{
    return regrid_core(1, n, x, y, w, sigma, overlap, x0, y0, 0.0, hnu, xo, yo, wo, io, jo);
}

void orc_free(void *p) { free(p); }

/* PopulationControl, src/python/blobs/base/regrid/_regridding.py:313-330:
 * repeat { flagged = |w| < thr; loss = sum |w[flagged]|; if loss < total:
 * accept; else thr /= 10 }; keep the unflagged blobs in order.  Writes the
 * kept indices into keep_idx (capacity n) and returns their count; the
 * threshold used is returned in *thr_used.  The loss is summed in index
 * order (numpy sums pairwise, so a loss within an ulp of the total
 * threshold may decide differently; parity tests avoid that edge).
 * Pinned by brute force in tests (direct evaluation of the loop).       */
int64_t orc_population_control(int64_t n, const double *w, double thr, double total,
                               int64_t *keep_idx, double *thr_used)
{
    for (;;) {
        double loss = 0.0;
        for (int64_t i = 0; i < n; i++)
            if (fabs(w[i]) < thr) loss += fabs(w[i]);
        if (loss < total || thr == 0.0) break;
        thr /= 10.0;
    }
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++)
        if (!(fabs(w[i]) < thr)) keep_idx[m++] = i;
    *thr_used = thr;
    return m;
}

/* ------------------------------------------------------------------------ */
/* A.5  Constant-strength straight vortex panels.
 * Influence matrix restates src/cpp/panels/cython/vortexPanel_Cython/
 * _vortexPanel_Cython_cpu.pyx:192-233 (inv_2pi = 1/(2*M_PI), inv_4pi =
 * inv_2pi/2 at :193-194).  Induced velocity restates :357-391.
 * Pinned by: brute-force Gauss-Legendre quadrature of point vortices along
 * each panel, and the cylinder potential-flow closed form
 * vx(0,1.5) = U(1 + R^2/y^2) of examples/panels/ex03 (:80).              */
void orc_panel_influence(int64_t M, const double *xCP, const double *yCP,
                         const double *xs, const double *ys,
                         const double *xe, const double *ye,
                         const double *cosA, const double *sinA, double *A)
{
    const double inv_2pi = 1.0 / (2.0 * M_PI);
    const double inv_4pi = inv_2pi / 2.0;
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < M; i++) {
        for (int64_t j = 0; j < M; j++) {
            double Lx = xe[j] - xs[j];
            double Ly = ye[j] - ys[j];
            double x2 = cosA[j] * Lx + sinA[j] * Ly;
            double xP = cosA[j] * (xCP[i] - xs[j]) + sinA[j] * (yCP[i] - ys[j]);
            double yP = -sinA[j] * (xCP[i] - xs[j]) + cosA[j] * (yCP[i] - ys[j]);
            double vxP = inv_2pi * (atan2(yP, (xP - x2)) - atan2(yP, xP));
            double vyP = -inv_4pi * log((xP * xP + yP * yP) /
                                        ((xP - x2) * (xP - x2) + yP * yP));
            double vx = cosA[j] * vxP - sinA[j] * vyP;
            double vy = sinA[j] * vxP + cosA[j] * vyP;
            A[i * M + j] = vx * cosA[i] + vy * sinA[i];
        }
    }
}

void orc_panel_induced(int64_t M, const double *gamma, int64_t ne,
                       const double *xE, const double *yE, const double *xs,
                       const double *ys, const double *xe, const double *ye,
                       const double *cosA, const double *sinA, double *vx,
                       double *vy)
{
    const double inv_2pi = 1.0 / (2.0 * M_PI);
    const double inv_4pi = inv_2pi / 2.0;
    int64_t i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < ne; i++) {
        double sx = 0.0, sy = 0.0;
        for (int64_t j = 0; j < M; j++) {
            double Lx = xe[j] - xs[j];
            double Ly = ye[j] - ys[j];
            double x2 = cosA[j] * Lx + sinA[j] * Ly;
            double xP = cosA[j] * (xE[i] - xs[j]) + sinA[j] * (yE[i] - ys[j]);
            double yP = -sinA[j] * (xE[i] - xs[j]) + cosA[j] * (yE[i] - ys[j]);
            double vxP = (gamma[j] * inv_2pi) * (atan2(yP, (xP - x2)) - atan2(yP, xP));
            double vyP = -(gamma[j] * inv_4pi) *
                         log((xP * xP + yP * yP) / ((xP - x2) * (xP - x2) + yP * yP));
            sx += cosA[j] * vxP - sinA[j] * vyP;
            sy += sinA[j] * vxP + cosA[j] * vyP;
        }
        vx[i] = sx;
        vy[i] = sy;
    }
}

/* ------------------------------------------------------------------------ */
/* FP64 LU with partial pivoting (row-major A, n x n), the algorithm of
 * LAPACK dgetrf that scipy.linalg.lu_factor calls at
 * src/python/panels/__panelsClass.py:961 (solve at :782).  Pivot = first
 * row of maximal |a| in the column.  Pinned against numpy.linalg.solve
 * (LAPACK) in tests.  Returns 0, or k+1 if U(k,k) == 0.                   */
int orc_lu_factor(int64_t n, double *A, int64_t *piv)
{
    int info = 0;
    for (int64_t k = 0; k < n; k++) {
        int64_t p = k;
        double best = fabs(A[k * n + k]);
        for (int64_t r = k + 1; r < n; r++) {
            double v = fabs(A[r * n + k]);
            if (v > best) { best = v; p = r; }
        }
        piv[k] = p;
        if (p != k)
            for (int64_t c = 0; c < n; c++) {
                double t = A[k * n + c];
                A[k * n + c] = A[p * n + c];
                A[p * n + c] = t;
            }
        double d = A[k * n + k];
        if (d == 0.0) {
            if (!info) info = (int)(k + 1);
            continue;
        }
        for (int64_t r = k + 1; r < n; r++) {
            double l = A[r * n + k] / d;
            A[r * n + k] = l;
            for (int64_t c = k + 1; c < n; c++) A[r * n + c] -= l * A[k * n + c];
        }
    }
    return info;
}

void orc_lu_solve(int64_t n, const double *LU, const int64_t *piv, double *b)
{
    for (int64_t k = 0; k < n; k++) {
        int64_t p = piv[k];
        if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int64_t r = 0; r < n; r++) {
        double s = b[r];
        for (int64_t c = 0; c < r; c++) s -= LU[r * n + c] * b[c];
        b[r] = s;
    }
    for (int64_t r = n - 1; r >= 0; r--) {
        double s = b[r];
        for (int64_t c = r + 1; c < n; c++) s -= LU[r * n + c] * b[c];
        b[r] = s / LU[r * n + r];
    }
}

/* ------------------------------------------------------------------------ */
/* Morton-key quad-tree binning and its stable sort.
 * The reference bins with an adaptive median-split tree (fmmsort.cpp:441-785);
 * that tree shape is not part of the contract.  This is the CPU statement of
 * the uniform Morton binning the GPU library uses, so the GPU keys,
 * permutation and leaf offsets can be compared bit for bit:
 *   ix = clamp(floor((x - x0) * inv_s), 0, 2^L - 1)   (FP64, no FMA)
 *   key = interleave(ix -> even bits, iy -> odd bits)
 *   perm = stable sort of keys (ties by input index).
 * Pinned by: sortedness, perm being a permutation, every point lying inside
 * the box its key names (brute force over boxes), tests.                  */
static uint32_t spread_bits(uint32_t v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

void orc_morton_keys(int64_t n, const double *x, const double *y, double x0,
                     double y0, double inv_s, int L, uint32_t *keys)
{
    const double hi = (double)((1u << L) - 1u);
    for (int64_t k = 0; k < n; k++) {
        double fx = floor((x[k] - x0) * inv_s);
        double fy = floor((y[k] - y0) * inv_s);
        if (!(fx >= 0.0)) fx = 0.0;
        if (fx > hi) fx = hi;
        if (!(fy >= 0.0)) fy = 0.0;
        if (fy > hi) fy = hi;
        keys[k] = spread_bits((uint32_t)fx) | (spread_bits((uint32_t)fy) << 1);
    }
}

typedef struct { uint32_t key; int32_t idx; } kv_t;
static int kv_cmp(const void *a, const void *b)
{
    const kv_t *p = (const kv_t *)a, *q = (const kv_t *)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    return p->idx < q->idx ? -1 : (p->idx > q->idx);
}

/* Stable sort of keys; writes sorted keys and the permutation
 * (sorted position -> input index).  Returns 0, -1 on allocation failure. */
int orc_stable_sort_keys(int64_t n, const uint32_t *keys, uint32_t *skeys,
                         int32_t *perm)
{
    kv_t *t = (kv_t *)malloc(sizeof(kv_t) * (size_t)(n > 0 ? n : 1));
    if (!t) return -1;
    for (int64_t k = 0; k < n; k++) { t[k].key = keys[k]; t[k].idx = (int32_t)k; }
    qsort(t, (size_t)n, sizeof(kv_t), kv_cmp);
    for (int64_t k = 0; k < n; k++) { skeys[k] = t[k].key; perm[k] = t[k].idx; }
    free(t);
    return 0;
}

/* Leaf offsets: start[k] = first sorted position with key >= k, for
 * k = 0..nleaf (start[nleaf] = n).                                        */
void orc_leaf_offsets(int64_t n, const uint32_t *skeys, int64_t nleaf,
                      int32_t *start)
{
    int64_t p = 0;
    for (int64_t k = 0; k <= nleaf; k++) {
        while (p < n && (int64_t)skeys[p] < k) p++;
        start[k] = (int32_t)p;
    }
}
