/*
 * phyflow_amd.h -- C ABI of libphyflow_amd, the MI355X (gfx950) vortex-blob
 * velocity hot path of pHyFlow.
 *
 * Conventions
 *   - Every function returns PHF_OK (0) or a negative PHF_ERR_* code; the
 *     message is available from phf_last_error() on the calling thread.
 *   - Array arguments are plain contiguous float64 host arrays unless the
 *     name says otherwise; the library never keeps a caller pointer after
 *     the call returns.  Outputs are written, never accumulated into.
 *   - Calls are thread-safe: each call borrows its own HIP stream and
 *     device workspace.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the pHyFlow source tree).
 */
#ifndef PHYFLOW_AMD_H
#define PHYFLOW_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PHF_OK 0
#define PHF_ERR_INVALID (-1)     /* bad argument: size, NULL, non-finite input */
#define PHF_ERR_DEVICE (-2)      /* HIP runtime error or no GPU */
#define PHF_ERR_NOMEM (-3)       /* device allocation failed or budget exceeded */
#define PHF_ERR_UNSUPPORTED (-4) /* option combination not provided */
#define PHF_ERR_BUSY (-5)        /* admission refused / queue full */
#define PHF_ERR_NOTFOUND (-6)    /* unknown handle or request id */

/* ---- precision of blob interactions ------------------------------------ */
#define PHF_PREC_F64 0  /* float64 throughout */
#define PHF_PREC_F32C 1 /* float32 pair arithmetic, compensated per-target sums */

/* ---- timed stages (HIP events on the call's streams) -------------------- */
/* The FMM runs its far field (P2M..L2L) on a second stream beside the near
 * field, so those stages overlap P2P in time. */
enum {
    PHF_STAGE_UPLOAD = 0, /* host -> device copies */
    PHF_STAGE_TREE,       /* bbox, Morton keys, radix sort, leaf offsets, gathers */
    PHF_STAGE_P2M,
    PHF_STAGE_M2M,
    PHF_STAGE_M2L,
    PHF_STAGE_L2L,
    PHF_STAGE_P2P,        /* near-field (or all-pairs) sum */
    PHF_STAGE_DOWNLOAD,   /* device -> host copies */
    PHF_STAGE_REMESH,
    PHF_STAGE_PANEL,
    PHF_STAGE_TREE_FAR,   /* tree through L2L replayed as captured graphs
                             (resident systems): TREE is then the tree graph
                             alone and P2M..L2L are -1 */
    PHF_STAGE_TOTAL,
    PHF_STAGE_COUNT
};

/* ---- library ------------------------------------------------------------- */

/* Message of the last failed call on this thread ("" if none). */
const char *phf_last_error(void);

/* Selects device 0 of this process and creates the stream pool.  Called
 * implicitly by every entry point; explicit calls are idempotent. */
int phf_init(void);

/* Device name, compute-unit count and HBM bytes of the selected GPU. */
int phf_device_info(char *name, int64_t name_len, int32_t *num_cu,
                    int64_t *hbm_bytes);

/* Waits until all work the library queued on the device has finished. */
int phf_synchronize(void);

/* Turn per-stage HIP-event timing on (1) or off (0, default). */
int phf_set_timing(int32_t enabled);

/* Per-stage milliseconds of the last timed call on this thread
 * (PHF_STAGE_COUNT doubles; -1 for a stage the call did not run). */
int phf_last_stage_times(double *ms);

/* ---- (a) direct Gaussian Biot-Savart ------------------------------------
 * Replaces vv2parGauss_gpu(xB,yB,wB,xT,yT,ksigmasqr,blocksize)
 *   src/cpp/blobs/cuda/vv2parGauss_gpu/vv2parGauss_py_gpu.cu:124-206
 * and its CPU twin vv2parGauss_C_cpu(xB,yB,wB,xT,yT,ksigmasqr)
 *   src/cpp/blobs/c/vv2parGauss_C/vv2parGauss_C_cpu.c:89-160.
 * u_t = inv_2pi * sum_m -dy*s, v_t = inv_2pi * sum_m dx*s,
 * s = w_m/r2*(1-exp(-r2/ksigmasqr)), r2 = dx^2+dy^2+1e-12.
 * No padding of nb/nt is required (the reference GPU path needed it). */
int phf_vv2par_gauss(const double *xb, const double *yb, const double *wb,
                     int64_t nb, const double *xt, const double *yt,
                     int64_t nt, double ksigmasqr, int32_t prec, double *vx,
                     double *vy);

/* Allocation-free variant of phf_vv2par_gauss.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same sum and outputs, but every array is DEVICE memory of
 * the library's GPU (each must lie inside one device allocation; a host
 * pointer is rejected with PHF_ERR_INVALID), and the call makes no device
 * or host allocation: the partial sums live in a file-scope device array of
 * 2^21 double2 (32 MiB) per GPU, held until the process ends.
 * Limit: nt <= 2,097,152 targets (larger nt is rejected with
 * PHF_ERR_INVALID before any launch).
 * stream: a hipStream_t of the caller; the work is enqueued there and the
 * call returns without waiting.  NULL: the library's own stream, and the
 * call returns when the results are written.
 * Every call on one GPU shares that GPU's scratch array, so calls that
 * overlap on one GPU (different streams or threads without ordering) race;
 * calls ordered on one stream do not.  phf_init() sets the array up; the
 * first call does it if phf_init() was not called. */
int phf_vv2par_gauss_noalloc(const double *xb, const double *yb,
                             const double *wb, int64_t nb, const double *xt,
                             const double *yt, int64_t nt, double ksigmasqr,
                             int32_t prec, double *vx, double *vy,
                             void *stream);

/* Algebraic cutoff kernel (kernel = 0 in velocity()/vorticity()).  Replace
 * vv2par_gpu / vv2par_C_cpu(xB,yB,wB,xT,yT,sigmasqr) and ww2par_gpu /
 * ww2par_C_cpu (src/cpp/blobs/c/vv2par_C/vv2par_C_cpu.c:89-153,
 * src/cpp/blobs/c/ww2par_C/ww2par_C_cpu.c:89-147, called with sigma^2 at
 * _induced.py:585, :695):
 *   u_t = inv_2pi sum_m -dy w_m/r2, v_t = inv_2pi sum_m dx w_m/r2,
 *   r2 = dx^2 + dy^2 + sigmasqr;  w_t = inv_2pi sum_m w_m / sqrt(r2). */
int phf_vv2par_cutoff(const double *xb, const double *yb, const double *wb,
                      int64_t nb, const double *xt, const double *yt,
                      int64_t nt, double sigmasqr, int32_t prec, double *vx,
                      double *vy);
int phf_ww2par_cutoff(const double *xb, const double *yb, const double *wb,
                      int64_t nb, const double *xt, const double *yt,
                      int64_t nt, double sigmasqr, int32_t prec, double *w);

/* Induced vorticity of Gaussian blobs.  Replaces
 * ww2parGauss_gpu(xB,yB,wB,xT,yT,ksigmasqr,blocksize)
 *   src/cpp/blobs/cuda/ww2parGauss_gpu/ww2parGauss_py_gpu.cu:104-200
 * and ww2parGauss_C_cpu(xB,yB,wB,xT,yT,ksigmasqr)
 *   src/cpp/blobs/c/ww2parGauss_C/ww2parGauss_C_cpu.c:89-153.
 * w_t = inv_pi * sum_m w_m exp(-(dx^2+dy^2+1e-12)/ksigmasqr) / ksigmasqr. */
int phf_ww2par_gauss(const double *xb, const double *yb, const double *wb,
                     int64_t nb, const double *xt, const double *yt,
                     int64_t nt, double ksigmasqr, int32_t prec, double *w);

/* ---- (b) 2-D FMM with the Oseen smoother --------------------------------
 * Replaces fmm2d_gpu / fmm2d_cpu(xB,yB,wB,xT,yT,Ndirect,xopt,cutoff,tol)
 *   src/cpp/blobs/c++/fmm2d_py_cpu_gpu/fmm2d_py.cpp:59-186 (GPU),
 *   :189-316 (CPU) -> fmmInteract fmm.cpp:88-304, directInteract
 *   direct.cpp:78.
 * Returns (qr, qi) = -sum_j w_j K(|z-z_j|)(conj(z-z_j)) exactly as the
 * reference (caller maps to velocity (-qi/2pi, -qr/2pi), _induced.py:334).
 * tol == 0 selects the direct sum over all pairs (direct.cpp:1409-1450).
 * opts may be NULL (defaults: order from tol by fmm.cpp:538-546, leaf level
 * from Ndirect by fmm.cpp:553, float64). */
typedef struct {
    int32_t order; /* expansion order p (coefficients 0..p); <=0: from tol */
    int32_t leaf;  /* target mean sources per leaf box; <=0: from Ndirect */
    int32_t prec;  /* PHF_PREC_F64 or PHF_PREC_F32C */
    int32_t ws;    /* well-separation distance in boxes (1 or 2); <=0: 1 */
} phf_fmm_opts;

int phf_fmm2d(const double *xb, const double *yb, const double *wb,
              int64_t nb, const double *xt, const double *yt, int64_t nt,
              int32_t ndirect, double xopt, double cutoff, double tol,
              const phf_fmm_opts *opts, double *qr, double *qi);

/* Debug/parity view of the Morton tree the FMM builds for nb points:
 * leaf level L, Morton keys in input order, the stable sort permutation
 * (sorted position -> input index) and leaf offsets start[0..4^L].
 * keys/perm hold nb entries, start holds 4^L+1 (pass NULL to skip).
 * x0,y0,inv_s receive the binning parameters (ix = floor((x-x0)*inv_s)). */
int phf_tree_build(const double *x, const double *y, int64_t n, int32_t L,
                   uint32_t *keys, int32_t *perm, int32_t *start, double *x0,
                   double *y0, double *inv_s);

/* ---- (c) M4' remeshing --------------------------------------------------
 * m4kernel_1d(x, c) -> w replaces
 *   src/cpp/blobs/c++/interpolation_kernels_1d/_interpolation_kernels_1d.cpp:21-93.
 * phf_regrid_m4 replaces Regrid(xBlob, yBlob, wBlob, sigma, overlap,
 *   xBounds, yBounds, interpKernel=0, c) of
 *   src/python/blobs/base/regrid/_regridding.py:36-222 (x0 = xBounds[0],
 *   y0 = yBounds[0]; h = sigma*overlap).  Output: every touched lattice cell
 *   (zero-weight ones included) in row-major (i, then j) order with
 *   x = i*h + h/2 + x0; integer cell indices are returned too.  Sums are
 *   float64, deterministic, without atomics.  The result is a library-owned
 *   blob set (its size is data-dependent). */
int phf_m4kernel_1d(const double *x, int64_t n, double c, double *w);

#define PHF_REMESH_M4 0    /* Regrid with M4' (c: diffusion parameter) */
#define PHF_REMESH_TUTTY 1 /* diffusion_tutty's weights */

typedef struct phf_blobset_s phf_blobset;
int phf_regrid_m4(const double *x, const double *y, const double *w, int64_t n, double sigma,
                  double overlap, double x0, double y0, double c, phf_blobset **out);
/* diffusion_tutty(dt, nu, h, xBlob, yBlob, wBlob, sigma, xBounds, yBounds,
 * overlap) of src/python/blobs/base/evolution/_evolution.py:238-445: the
 * remesh above (h = sigma*overlap, :333; same cells, order and coordinates)
 * with Tutty's diffusion weights, hNu = sqrt(dt*nu) (:366, :385-405). */
int phf_diffusion_tutty(const double *x, const double *y, const double *w, int64_t n,
                        double sigma, double overlap, double x0, double y0, double dt, double nu,
                        phf_blobset **out);
int64_t phf_blobset_size(const phf_blobset *b);
/* Copies the blob set into caller arrays of phf_blobset_size() entries
 * (any pointer may be NULL). */
int phf_blobset_copy(const phf_blobset *b, double *x, double *y, double *w, int64_t *i,
                     int64_t *j);
void phf_blobset_free(phf_blobset *b);

/* Replaces PopulationControl(xBlob, yBlob, wBlob, wThreshold,
 * wTotalThreshold) (_regridding.py:227-330): blobs with |w| < threshold are
 * dropped; the threshold starts at wThreshold and is divided by 10 until
 * the dropped total |w| is below wTotalThreshold; order is kept. */
int phf_population_control(const double *x, const double *y, const double *w, int64_t n,
                           double wThreshold, double wTotalThreshold, phf_blobset **out);

/* ---- (d) vortex panels and the dense panel solve ---------------------------
 * phf_panel_influence replaces assemble_influenceMatrix(xCP, yCP, xPS, yPS,
 *   xPE, yPE, cosA, sinA) -> A[M,M] (row-major),
 *   src/cpp/panels/cython/vortexPanel_Cython/_vortexPanel_Cython_cpu.pyx:81-233.
 * phf_panel_induced_velocity replaces inducedVelocity(gamma, xEval, yEval,
 *   xPS, yPS, xPE, yPE, cosA, sinA) -> (vx, vy), same file :243-391.
 * phf_dense_solve replaces the solves of Panels.solve
 *   (src/python/panels/__panelsClass.py:715-822): rows == n is the square
 *   'fixed' system (scipy lu_factor/lu_solve, :961/:782); rows > n is the
 *   'prescribed' least-squares system [A; L] x = [RHS; -gTotal]
 *   (numpy.linalg.lstsq, :800-822).  B is row-major rows x n.  Float32 LU
 *   with partial pivoting (MFMA trailing updates) + float64 iterative
 *   refinement (normal equations with semi-normal float64 residuals when
 *   rows > n).  info may be NULL. */
int phf_panel_influence(const double *xCP, const double *yCP, const double *xs,
                        const double *ys, const double *xe, const double *ye,
                        const double *cosA, const double *sinA, int64_t m, double *A);
int phf_panel_induced_velocity(const double *gamma, int64_t m, const double *xE,
                               const double *yE, int64_t ne, const double *xs,
                               const double *ys, const double *xe, const double *ye,
                               const double *cosA, const double *sinA, double *vx, double *vy);

typedef struct {
    int32_t iterations; /* refinement sweeps */
    int32_t converged;  /* 1: correction reached float64 resolution */
    double lu_ms;       /* device time of the float32 factorisation */
    double refine_ms;   /* device time of the refinement */
    double total_ms;    /* wall time of the call */
} phf_solve_info;

/* The factorisation behind phf_dense_solve on its own: float32 LU with
 * partial pivoting of the column-major n x n A, in LAPACK getrf form (the
 * factor the reference takes through scipy.linalg.lu_factor,
 * __panelsClass.py:961): LU holds U and the unit-lower L, ipiv[k] (0-based)
 * the row exchanged with row k at step k, info = k + 1 for the first zero
 * pivot U[k][k] (0 if none). */
int phf_lu_factor_f32(const float *A, int64_t n, float *LU, int32_t *ipiv, int32_t *info);

/* Float32 Cholesky A = L Lᵀ of the symmetric positive definite
 * column-major n x n A (n <= 1024), the factorisation phf_dense_solve uses
 * for the normal matrix BᵀB of a least-squares system (in place of the
 * reference's numpy.linalg.lstsq, __panelsClass.py:822), returned in the
 * getrf form of phf_lu_factor_f32 with identity pivots: LU holds the unit
 * lower L D^-1 and the upper D Lᵀ, D = diag(L).  info = k + 1 for the first
 * pivot <= 0 (not positive definite in float32; 0 if none). */
int phf_chol_factor_f32(const float *A, int64_t n, float *LU, int32_t *info);

int phf_dense_solve(const double *B, int64_t rows, int64_t n, const double *c, double *x,
                    phf_solve_info *info);

/* ---- resident blob systems -----------------------------------------------
 * A phf_system keeps one blob set (x, y, w) resident in HBM with its own HIP
 * stream and workspace, so repeated evaluations make no host copies and no
 * device allocations.  It stands in for the state the reference's Blobs
 * object holds on the host (src/python/blobs/__blobsClass.py:415-494,
 * properties x, y, g at :2651-2690) and re-uploads on every GPU call
 * (cudaeval.cu:154-342). */
typedef struct phf_system_s phf_system;

#define PHF_METHOD_FMM 0    /* blobOptions.FMM_METHOD */
#define PHF_METHOD_DIRECT 1 /* blobOptions.DIRECT_METHOD */

/* Velocity-evaluation parameters, the arguments of the reference's
 * velocity(..., sigma, k, ..., method, Ndirect, tol, cutoff)
 * (src/python/blobs/base/induced/_induced.py:34-36).  method FMM uses the
 * fmm2d Oseen smoother (tol == 0: its direct sum); method DIRECT uses the
 * vv2parGauss EPS-regularised Gaussian kernel. */
typedef struct {
    int32_t method;  /* PHF_METHOD_FMM or PHF_METHOD_DIRECT */
    int32_t prec;    /* PHF_PREC_F64 or PHF_PREC_F32C */
    int32_t order;   /* FMM order; <=0: from tol (fmm.cpp:538-546) */
    int32_t leaf;    /* FMM mean sources per leaf; <=0: from ndirect */
    int32_t ws;      /* FMM well-separation in boxes; <=0: 1 */
    int32_t ndirect; /* reference Ndirect (default 35) */
    double sigma;    /* blob core size */
    double k;        /* core multiplier, k*sigma^2 (default 2) */
    double tol;      /* FMM tolerance (default 1e-6); 0: direct Oseen sum */
    double cutoff;   /* Oseen cutoff; <=0: 5*xopt (_induced.py:240) */
    double leaf_side; /* FMM domain: 0 cost model, < 0 the bounding box, > 0 leaves at
                         least leaf_side * cutoff wide (set by the autotuner) */
} phf_vel_params;

/* Work and timing of the last evaluation (stage ms are HIP-event times on
 * the system's stream; -1 for stages not run).  Set flags = 1 on input to
 * also count the FMM work (p2p_pairs, m2l_links; adds small kernels).
 * A resident system's FMM records its launch sequence (tree through L2L)
 * as a HIP graph and replays it while the point count, parameters and the
 * domain side stay the same; flags |= PHF_INFO_EAGER launches every
 * kernel one by one instead (same results).  flags |= PHF_INFO_TIME_P2P
 * times only the near-field (P2P) stage, with two events on the stream
 * (the other stages report -1). */
#define PHF_INFO_COUNT_WORK 1
#define PHF_INFO_EAGER 2
#define PHF_INFO_TIME_P2P 4
typedef struct {
    int32_t L, p, nnear, flags;
    double D;
    int64_t p2p_pairs; /* FMM: near-field pairs; direct: nb*nt */
    int64_t m2l_links;
    double ms[PHF_STAGE_COUNT];
} phf_eval_info;

int phf_system_create(const double *x, const double *y, const double *w, int64_t n,
                      phf_system **out);
int phf_system_destroy(phf_system *s);
int phf_system_size(phf_system *s, int64_t *n);
int phf_system_get(phf_system *s, double *x, double *y, double *w);
int phf_system_set(phf_system *s, const double *x, const double *y, const double *w,
                   int64_t n);

/* Velocity at the blob centres, the reference's
 * velocity(xBlob, yBlob, wBlob, sigma, k, kernel=gauss, xEval=None, ...)
 * (_induced.py:34-166) with (vx, vy) in its convention.  vx/vy may be NULL:
 * the result then stays on the device (phf_system_get_velocity fetches
 * it).  info may be NULL. */
int phf_system_velocity(phf_system *s, const phf_vel_params *p, double *vx, double *vy,
                        phf_eval_info *info);
int phf_system_get_velocity(phf_system *s, double *vx, double *vy);

/* ---- resident panel bodies and the coupled step -----------------------------
 * A phf_panels object holds one panel discretisation (start/end points of m
 * straight vortex panels), its geometry (__panelsClass.py:1036-1058: panel
 * lengths, angles, collocation points CP = mid - 0.1 L n), its influence
 * matrix and the float32 LU factors of A ('fixed') or of the least-squares
 * system [A; L] ('prescribed', Kelvin row sum_j L_j s_j = -gTotal) -- the
 * state of the reference's Panels object (__panelsClass.py:446-525). */
typedef struct phf_panels_s phf_panels;
#define PHF_PANELS_FIXED 0
#define PHF_PANELS_PRESCRIBED 1
#define PHF_PANELS_MOVING 2  /* square system A s = RHS, rebuilt by phf_panels_update */

int phf_panels_create(const double *xs, const double *ys, const double *xe, const double *ye,
                      int64_t m, int32_t problem, double gTotal, phf_panels **out);
int phf_panels_destroy(phf_panels *p);
/* m, collocation points (m each, may be NULL) and A (m*m row-major, may be NULL). */
int phf_panels_info(phf_panels *p, int64_t *m, double *xcp, double *ycp, double *A);
/* Panels.solve(vxExternal, vyExternal) (__panelsClass.py:715-822) with the
 * cached factors: strengths for RHS = -(v . t). */
int phf_panels_solve(phf_panels *p, const double *vx, const double *vy, double *strength,
                     int32_t *iterations);

/* Device times (ms) of one rebuild of a panel body. */
typedef struct {
    double influence_ms; /* influence matrix (+ Kelvin row) */
    double normal_ms;    /* float64 copies, B^T B for 'prescribed' */
    double lu_ms;        /* float32 LU with partial pivoting (matrix cores) */
    double inverse_ms;   /* float32 inverse from the factors */
    double total_ms;     /* all of the above + grouped-velocity factors */
} phf_panels_timing;

/* Panels.updateBody (__panelsClass.py:530-638): new start/end points of
 * the same m panels (a moving or deforming body).  Geometry, collocation
 * points, the influence matrix (__assembleInfluenceMatrix :872-961) and the
 * factors are rebuilt on the device; timing may be NULL. */
int phf_panels_update(phf_panels *p, const double *xs, const double *ys, const double *xe,
                      const double *ye, phf_panels_timing *timing);

/* The same rebuild, enqueued on the body's stream without waiting for it
 * (no counterpart in the reference, whose updateBody returns with the
 * factors done).  The arrays are copied before the call returns.  The next
 * call on the body completes the rebuild first and reports its errors (a
 * singular matrix): phf_panels_solve, phf_panels_info, phf_panels_update,
 * phf_panels_update_async, phf_system_step_rk2.  In the coupled step only
 * the panel solve waits for the factors, so the first blob velocity
 * evaluation runs beside the factorisation. */
int phf_panels_update_async(phf_panels *p, const double *xs, const double *ys, const double *xe,
                            const double *ye);

typedef struct {
    double dt;
    double vinf_x, vinf_y;   /* free stream */
    int32_t panel_update;    /* 0: solve panels at stage 1 ('constant'), 1: every stage */
    int32_t remesh;          /* 1: remesh after convection */
    double remesh_sigma, remesh_overlap, remesh_x0, remesh_y0;
    double remesh_c;         /* M4': the diffusion parameter c (Wee); Tutty: hNu = sqrt(nu dt) */
    int32_t pc_enabled;      /* 1: population control after the remesh */
    int32_t remesh_kernel;   /* PHF_REMESH_M4 or PHF_REMESH_TUTTY */
    double pc_w_threshold, pc_total_threshold; /* gThresholdLocal, gThresholdGlobal */
} phf_step_params;

typedef struct {
    int64_t n_before, n_after;
    /* float64 invariants before/after the remesh: total circulation, linear
     * impulse (sum w y, -sum w x) and sum w (x^2 + y^2) */
    double circ_before, circ_after;
    double impulse_x_before, impulse_x_after, impulse_y_before, impulse_y_after;
    double m2_before, m2_after;
    int32_t solve_iterations, pad;
    /* convection and remesh: elapsed time between events on the step's
     * stream; total: host wall time of the call */
    double ms_convection, ms_remesh, ms_total;
} phf_step_info;

/* One time step of the resident blobs: Heun RK2 convection with blob
 * velocities (params vp), panel velocities (panels may be NULL) and the
 * free stream, then an optional M4' remesh.  Replaces the convection and
 * redistribution of LagrangianSolver.evolve / Blobs.evolve
 * (__lagrangianSolverClass.py:269-375, __blobsClass.py:705-781), whose
 * integrator is RK4; RK2 here.  info may be NULL. */
int phf_system_step_rk2(phf_system *s, const phf_vel_params *vp, phf_panels *panels,
                        const phf_step_params *sp, phf_step_info *info);

/* ---- batched systems ---------------------------------------------------------
 * Many independent blob systems of uneven sizes (e.g. a bluff-body
 * parameter sweep: one Blobs object per case in the reference, each calling
 * fmm2d on its own, fmm2d_py.cpp:59) evaluated together: every FMM phase is
 * one launch over all systems through a segment table, and the evaluation
 * is captured once as a hipGraph and replayed.  Velocities at the blob
 * centres in the velocity() convention (_induced.py:334-336); each system
 * has its own core size sigma; k, order, leaf size, well-separation and
 * precision are shared. */
typedef struct phf_batch_s phf_batch;
typedef struct {
    int32_t prec;   /* PHF_PREC_F64 or PHF_PREC_F32C */
    int32_t order;  /* FMM order (<= 0: 16) */
    int32_t leaf;   /* mean sources per leaf (<= 0: 64) */
    int32_t ws;     /* well-separation (<= 0: 1) */
    double k;       /* core multiplier (k sigma^2) */
} phf_batch_params;

typedef struct {
    double ms;            /* device time of the evaluation (HIP events) */
    int32_t graph;        /* 1 if replayed from the captured graph */
    int32_t flags;        /* nonzero: a near set overflowed (results invalid) */
    int64_t total_points;
    int32_t levels_max, pad;
    int64_t remeshed;     /* phf_batch_step: blobs after the remesh, all systems */
    int32_t spilled;      /* phf_batch_step: systems remeshed by the single-system
                             path because their lattice range exceeded the
                             reserved cells */
    int32_t pad2;
} phf_batch_info;

/* Convection + remesh of every system after the batched velocity
 * (phf_batch_step): x <- x + dt (vx + vinf_x) (_evolution.py:540-541), then
 * Regrid (_regridding.py:118-218) onto the lattice of spacing
 * h_s = sigma_s * overlap with origin (x0h h_s, y0h h_s) (Blobs.redistribute:
 * x0h = y0h = -1.5, __blobsClass.py:891-897) and M4' parameter c. */
typedef struct {
    double dt;            /* 0: remesh the current positions */
    double vinf_x, vinf_y;
    double overlap;
    double x0h, y0h;
    double c;
    double cell_factor;   /* lattice cells reserved per system:
                             cell_factor n_s + 16 sqrt(n_s) + 256 (<= 0: 2) */
} phf_batch_remesh_params;

int phf_batch_create(const int64_t *sizes, const double *sigma, int32_t count,
                     const phf_batch_params *bp, phf_batch **out);
int phf_batch_destroy(phf_batch *b);
int phf_batch_set_system(phf_batch *b, int32_t s, const double *x, const double *y,
                         const double *w);
/* Evaluate every system; use_graph = 1 captures the launch sequence on the
 * first call and replays it afterwards.  info may be NULL. */
int phf_batch_velocity(phf_batch *b, int32_t use_graph, phf_batch_info *info);
int phf_batch_get_velocity(phf_batch *b, int32_t s, double *vx, double *vy);
/* Enable phf_batch_step with these parameters (reserves the lattice
 * workspace; drops a previously captured step graph). */
int phf_batch_set_remesh(phf_batch *b, const phf_batch_remesh_params *rp);
/* Velocity of every system, then convection + remesh, as one launch
 * sequence (use_graph = 1: captured on first use, replayed after).  The
 * systems' own blobs are not changed; the remeshed blobs are read with
 * phf_batch_get_remeshed. */
int phf_batch_step(phf_batch *b, int32_t use_graph, phf_batch_info *info);
/* Remeshed blobs of system s from the last phf_batch_step, row-major lattice
 * order; *m receives the count; x, y, w (capacity *m) may be NULL to query
 * the count only. */
int phf_batch_get_remeshed(phf_batch *b, int32_t s, int64_t *m, double *x, double *y, double *w);

/* ---- serving --------------------------------------------------------------
 * A request loop over resident simulations (no counterpart in the
 * reference, whose callers step their Blobs synchronously).  Simulations
 * are admitted by HBM footprint; step requests (one RK2 step + optional
 * remesh each, phf_system_step_rk2 without panels) are queued on arrival
 * and run by `slots` workers as slots free up, each simulation on its own
 * HIP stream so concurrent steps overlap on the device.  Latency is
 * measured from submission to completion. */
typedef struct phf_server_s phf_server;

typedef struct {
    int32_t status, pad; /* 0 done, < 0 error code */
    double latency_ms;   /* submission -> completion */
    double queue_ms;     /* submission -> start */
    phf_step_info info;
} phf_request_result;

typedef struct {
    int64_t completed, failed;
    double p50_ms, p99_ms, max_ms; /* latency percentiles of completed requests */
    double span_s;                 /* first submission -> last completion */
    int64_t hbm_used, hbm_budget;
    int32_t slots, pad;
} phf_serve_stats;

/* hbm_budget_bytes <= 0: 7/8 of the device memory. */
int phf_server_create(int32_t slots, int64_t hbm_budget_bytes, phf_server **out);
int phf_server_destroy(phf_server *s);
/* PHF_ERR_BUSY if the simulation's footprint does not fit the budget.  The
 * caller keeps ownership of sys and must not use it while registered. */
int phf_server_add_system(phf_server *s, phf_system *sys, int64_t *id);
int phf_server_remove_system(phf_server *s, int64_t id);
int phf_server_submit(phf_server *s, int64_t sim, const phf_vel_params *vp,
                      const phf_step_params *sp, int64_t *req_id);
/* timeout_s < 0: wait indefinitely.  The request is forgotten once waited. */
int phf_server_wait(phf_server *s, int64_t req_id, double timeout_s, phf_request_result *res);
int phf_server_drain(phf_server *s);
int phf_server_stats(phf_server *s, int32_t reset, phf_serve_stats *st);

#ifdef __cplusplus
}
#endif

#endif /* PHYFLOW_AMD_H */
