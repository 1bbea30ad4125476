// phf_step.hip -- resident vortex-panel bodies and one coupled time step
// (convection by RK2 with FMM/direct blob velocities, panel velocities and
// the free stream, then optional M4' remesh with float64 invariants).
//
// Reference call pattern (restated, not copied):
//   LagrangianSolver.__coupled_convection, __lagrangianSolverClass.py:496-677:
//   per stage the blobs' velocity at the panel collocation points plus vInf
//   gives the slip velocity, Panels.solve() the panel strengths (stage 1
//   only for the default 'constant' update, every stage for 'varying'),
//   then blob velocity = blobs->blobs + panels->blobs + vInf.
//   Blobs.redistribute (__blobsClass.py:891-897) remeshes with bounds
//   -1.5 h.
// The reference integrates with RK4 and has no RK2; the step here is Heun's
// second-order scheme, x1 = x0 + dt k(x0), x = x0 + dt/2 (k(x0) + k(x1)),
// with every stage evaluated at its own positions (the reference RK4 reuses
// aliased positions, a defect not reproduced here).
#include "phf_panel.hpp"
#include "phf_remesh.hpp"
#include "phf_system.hpp"

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace phf {

struct Panels {
    std::mutex mu;
    Lane lane;  // arena holds the persistent panel data (never reset)
    int64_t m = 0;
    int problem = 0;  // PHF_PANELS_FIXED, _PRESCRIBED or _MOVING
    double gTotal = 0.0;
    double *xs = nullptr, *ys = nullptr, *xe = nullptr, *ye = nullptr;
    double *cosA = nullptr, *sinA = nullptr, *L = nullptr, *xcp = nullptr, *ycp = nullptr;
    double *A = nullptr;  // influence matrix, row-major m x m
    DenseFactors f;       // factors of A or of [A; L]
    double *s = nullptr;  // last strengths
    PanelSeriesGeom sg;   // grouped-velocity factors of this body
    std::vector<double> hxcp, hycp;
    int last_iterations = 0;
    bool have_solution = false;  // s holds a converged solution
    // A rebuild enqueued on the body's stream and not yet completed
    // (phf_panels_update_async): ev_geom follows the copies of the new
    // geometry, ev_done the whole rebuild, after which info_host holds the
    // factorisation's info.  geom_pin stages the geometry (9 m doubles) so
    // the copies do not read the caller's arrays after the call returns.
    bool pending = false;
    hipEvent_t ev_geom = nullptr, ev_done = nullptr;
    double *geom_pin = nullptr;
    int32_t *info_host = nullptr;
};

static int panels_rebuild_finish(Panels *P, phf_panels_timing *timing = nullptr);

// ------------------------------------------------------------- kernels ----
// RHS of the no-slip condition, -(v . t) (__panelsClass.py:771); for the
// 'prescribed' problem the last entry is -gTotal.
__global__ void slip_rhs_kernel(const double *__restrict__ vx, const double *__restrict__ vy,
                                int64_t m, double vinf_x, double vinf_y,
                                const double *__restrict__ cosA, const double *__restrict__ sinA,
                                int prescribed, double gTotal, double *__restrict__ rhs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) rhs[i] = (-(vx[i] + vinf_x)) * cosA[i] + (-(vy[i] + vinf_y)) * sinA[i];
    if (prescribed && i == m) rhs[m] = -gTotal;
}

// k = v_blobs + v_panels + vInf
__global__ void total_velocity(double *__restrict__ kx, double *__restrict__ ky,
                               const double *__restrict__ px, const double *__restrict__ py,
                               int64_t n, double vinf_x, double vinf_y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    kx[i] = (kx[i] + (px ? px[i] : 0.0)) + vinf_x;
    ky[i] = (ky[i] + (py ? py[i] : 0.0)) + vinf_y;
}

// (kx, ky) = (vx + px + vinf_x, vy + py + vinf_y) for the first n entries
__global__ void total_velocity_to(const double *__restrict__ vx, const double *__restrict__ vy,
                                  const double *__restrict__ px, const double *__restrict__ py,
                                  int64_t n, double vinf_x, double vinf_y,
                                  double *__restrict__ kx, double *__restrict__ ky)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    kx[i] = (vx[i] + px[i]) + vinf_x;
    ky[i] = (vy[i] + py[i]) + vinf_y;
}

// FMM points of a coupled stage: the n blobs (x, y, w), then the m
// collocation points with zero circulation
__global__ void concat_points(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ w, int64_t n,
                              const double *__restrict__ xc, const double *__restrict__ yc,
                              int64_t m, double *__restrict__ tx, double *__restrict__ ty,
                              double *__restrict__ tw)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        tx[i] = x[i];
        ty[i] = y[i];
        tw[i] = w[i];
    } else if (i < n + m) {
        tx[i] = xc[i - n];
        ty[i] = yc[i - n];
        tw[i] = 0.0;
    }
}

__global__ void euler_stage(const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ kx, const double *__restrict__ ky,
                            int64_t n, double dt, double *__restrict__ x1,
                            double *__restrict__ y1)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x1[i] = x[i] + dt * kx[i];
    y1[i] = y[i] + dt * ky[i];
}

// the remeshed blobs into the system's arrays: one launch for the three
__global__ void copy3_f64(const double *__restrict__ ax, const double *__restrict__ ay,
                          const double *__restrict__ aw, int64_t n, double *__restrict__ bx,
                          double *__restrict__ by, double *__restrict__ bw)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        bx[i] = ax[i];
        by[i] = ay[i];
        bw[i] = aw[i];
    }
}

__global__ void heun_final(double *__restrict__ x, double *__restrict__ y,
                           const double *__restrict__ k1x, const double *__restrict__ k1y,
                           const double *__restrict__ k2x, const double *__restrict__ k2y,
                           int64_t n, double half_dt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = x[i] + half_dt * (k1x[i] + k2x[i]);
    y[i] = y[i] + half_dt * (k1y[i] + k2y[i]);
}

// ------------------------------------------------------------ helpers -----
static int panels_solve_device(Lane *L, Panels *P, const double *vx_cp, const double *vy_cp,
                               double vinf_x, double vinf_y)
{
    const int64_t rows = P->f.rows;
    double *rhs;
    PHF_ALLOC(rhs, L->arena, double, (size_t)rows);
    slip_rhs_kernel<<<div_up(rows, 256), 256, 0, L->stream>>>(
        vx_cp, vy_cp, P->m, vinf_x, vinf_y, P->cosA, P->sinA,
        P->problem == PHF_PANELS_PRESCRIBED ? 1 : 0, P->gTotal, rhs);
    PHF_HIP(hipGetLastError());
    // start from the last strengths (the slip condition changes little from
    // one stage to the next) with a first burst of one sweep more than the
    // last solve took
    SolveStats st;
    const bool warm = P->have_solution;
    P->have_solution = false;
    PHF_TRY(dense_refine_device(L, P->f, rhs, P->s, &st, warm,
                                warm ? std::min(6, P->last_iterations + 1) : 6));
    P->last_iterations = st.iterations;
    P->have_solution = st.converged;
    if (!st.converged) {
        set_error("panel solve did not converge (%d refinement sweeps)", st.iterations);
        return PHF_ERR_INVALID;
    }
    return PHF_OK;
}

// Total velocity k at positions (X, Y) of the n blobs carrying w.
static int stage_velocity(Lane *L, const double *X, const double *Y, const double *w, int64_t n,
                          const phf_vel_params &vp, Panels *P, bool solve_panels, double vinf_x,
                          double vinf_y, double *kx, double *ky, StageTimer &tm, bool *used)
{
    FmmInfo fi;
    if (!P) {
        PHF_TRY(velocity_device(L, X, Y, w, n, X, Y, n, true, vp, kx, ky, tm, used, &fi));
        total_velocity<<<div_up(n, 256), 256, 0, L->stream>>>(kx, ky, nullptr, nullptr, n,
                                                             vinf_x, vinf_y);
        PHF_HIP(hipGetLastError());
        return PHF_OK;
    }
    const int64_t m = P->m, nt = n + m;
    double *tx, *ty, *tw, *vx, *vy, *px, *py;
    PHF_ALLOC(tx, L->arena, double, (size_t)nt);
    PHF_ALLOC(ty, L->arena, double, (size_t)nt);
    PHF_ALLOC(tw, L->arena, double, (size_t)nt);
    PHF_ALLOC(vx, L->arena, double, (size_t)nt);
    PHF_ALLOC(vy, L->arena, double, (size_t)nt);
    PHF_ALLOC(px, L->arena, double, (size_t)std::max<int64_t>(nt, 1));
    PHF_ALLOC(py, L->arena, double, (size_t)std::max<int64_t>(nt, 1));
    // a rebuild still in flight on the body's stream: the collocation points
    // are needed now, the factors only for the solve below, so the blob
    // velocity evaluation runs beside the factorisation
    if (P->pending) PHF_HIP(hipStreamWaitEvent(L->stream, P->ev_geom, 0));
    concat_points<<<div_up(nt, 256), 256, 0, L->stream>>>(X, Y, w, n, P->xcp, P->ycp, m, tx, ty,
                                                          tw);
    PHF_HIP(hipGetLastError());
    // blobs -> blobs and blobs -> collocation points in one evaluation: the
    // collocation points join the sources with zero circulation (exact
    // zeros in every sum), so targets and sources share one tree
    PHF_TRY(velocity_device(L, tx, ty, tw, nt, tx, ty, nt, true, vp, vx, vy, tm, used, &fi));
    PHF_TRY(panels_rebuild_finish(P));
    if (solve_panels) {
        tm.begin(PHF_STAGE_PANEL);
        used[PHF_STAGE_PANEL] = true;
        PHF_TRY(panels_solve_device(L, P, vx + n, vy + n, vinf_x, vinf_y));
        tm.end(PHF_STAGE_PANEL);
    }
    // the panel velocity at the same points in the tree's Morton order when
    // the evaluation left it (the collocation points' values go unused),
    // else at the blobs, sorted by panel_induced_device itself
    const int32_t *ord = fi.sorted_pay ? reinterpret_cast<const int32_t *>(fi.sorted_pay) + 3
                                       : fi.sorted_perm;
    if (ord && fi.sorted_n == nt)
        PHF_TRY(panel_induced_device(L, m, P->s, nt, tx, ty, P->xs, P->ys, P->xe, P->ye,
                                     P->cosA, P->sinA, px, py, vp.prec, &P->sg, ord,
                                     fi.sorted_pay ? 4 : 1));
    else
        PHF_TRY(panel_induced_device(L, m, P->s, n, X, Y, P->xs, P->ys, P->xe, P->ye, P->cosA,
                                     P->sinA, px, py, vp.prec, &P->sg));
    total_velocity_to<<<div_up(n, 256), 256, 0, L->stream>>>(vx, vy, px, py, n, vinf_x, vinf_y,
                                                             kx, ky);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

static void moments_to(const double4 &m, double *circ, double *ix, double *iy, double *m2)
{
    // linear impulse I = (sum w y, -sum w x)
    *circ = m.x;
    *ix = m.z;
    *iy = -m.y;
    *m2 = m.w;
}

// Panel geometry on the host (__panelsClass.py:1036-1058): L = |end - start|,
// (cos, sin) = (end - start) / L, n = (-sin, cos), CP = mid - n L 0.1.
struct PanelGeomHost {
    const double *xs, *ys, *xe, *ye;
    std::vector<double> cs, sn, Ln, xc, yc;
};

static int panel_geometry_host(const double *xs, const double *ys, const double *xe,
                               const double *ye, int64_t m, PanelGeomHost *g)
{
    PHF_TRY(check_host_array(xs, m, "xPanelStart"));
    PHF_TRY(check_host_array(ys, m, "yPanelStart"));
    PHF_TRY(check_host_array(xe, m, "xPanelEnd"));
    PHF_TRY(check_host_array(ye, m, "yPanelEnd"));
    g->xs = xs;
    g->ys = ys;
    g->xe = xe;
    g->ye = ye;
    g->cs.resize(m);
    g->sn.resize(m);
    g->Ln.resize(m);
    g->xc.resize(m);
    g->yc.resize(m);
    for (int64_t j = 0; j < m; ++j) {
        const double lx = xe[j] - xs[j], ly = ye[j] - ys[j];
        g->Ln[j] = std::sqrt(lx * lx + ly * ly);
        PHF_REQUIRE(g->Ln[j] > 0.0, "panel %lld has zero length", (long long)j);
        g->cs[j] = lx / g->Ln[j];
        g->sn[j] = ly / g->Ln[j];
        const double nx = -g->sn[j], ny = g->cs[j];
        g->xc[j] = 0.5 * (xe[j] + xs[j]) - nx * g->Ln[j] * 0.1;
        g->yc[j] = 0.5 * (ye[j] + ys[j]) - ny * g->Ln[j] * 0.1;
    }
    return PHF_OK;
}

__global__ void zero_doubles(double *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// New geometry into the body's device buffers, then its influence matrix
// (with the Kelvin row for 'prescribed'), the float32 factors and the
// grouped-velocity factors, all rebuilt in place (__panelsClass.py:
// updateBody :530-638 -> __assembleInfluenceMatrix :872-961), enqueued on
// the body's stream without a host synchronisation.  Any earlier rebuild
// must be finished (the staging buffer and the events are reused).
static int panels_rebuild_enqueue(Panels *P, const PanelGeomHost &g)
{
    Lane *L = &P->lane;
    const int64_t m = P->m;
    const double *host[9] = {g.xs, g.ys, g.xe, g.ye, g.cs.data(), g.sn.data(), g.Ln.data(),
                             g.xc.data(), g.yc.data()};
    double *dev[9] = {P->xs, P->ys, P->xe, P->ye, P->cosA, P->sinA, P->L, P->xcp, P->ycp};
    int rc = [&]() -> int {
        for (int k = 0; k < 9; ++k) {
            std::memcpy(P->geom_pin + k * m, host[k], sizeof(double) * m);
            PHF_HIP(hipMemcpyAsync(dev[k], P->geom_pin + k * m, sizeof(double) * m,
                                   hipMemcpyHostToDevice, L->stream));
        }
        PHF_HIP(hipEventRecord(P->ev_geom, L->stream));
        P->hxcp = g.xc;
        P->hycp = g.yc;
        hipEvent_t *ev = L->ev;  // [0] start, [1] influence, [2..4] factor marks
        PHF_HIP(hipEventRecord(ev[0], L->stream));
        PHF_TRY(panel_influence_device(L, m, P->xcp, P->ycp, P->xs, P->ys, P->xe, P->ye,
                                       P->cosA, P->sinA, P->A));
        if (P->problem == PHF_PANELS_PRESCRIBED)  // Kelvin row: sum_j L_j s_j = -gTotal (:509-518)
            PHF_HIP(hipMemcpyAsync(P->A + m * m, P->L, sizeof(double) * m,
                                   hipMemcpyDeviceToDevice, L->stream));
        PHF_HIP(hipEventRecord(ev[1], L->stream));
        const int64_t rows = P->problem == PHF_PANELS_PRESCRIBED ? m + 1 : m;
        PHF_TRY(dense_factor_enqueue(L, P->A, rows, m, &P->f, ev + 2));
        PHF_TRY(panel_series_geom_device(L, m, P->xs, P->ys, P->xe, P->ye, P->cosA, P->sinA,
                                         &P->sg));
        PHF_HIP(hipEventRecord(ev[5], L->stream));
        PHF_HIP(hipMemcpyAsync(P->info_host, P->f.info, sizeof(int32_t), hipMemcpyDeviceToHost,
                               L->stream));
        PHF_HIP(hipEventRecord(P->ev_done, L->stream));
        return PHF_OK;
    }();
    if (rc != PHF_OK) {
        (void)hipStreamSynchronize(L->stream);  // nothing left in flight
        return rc;
    }
    P->pending = true;
    return PHF_OK;
}

// Complete an enqueued rebuild: wait for it, take the pivoted LU if the
// Cholesky met a non-positive pivot, and report a singular matrix.  timing
// (optional): device times of the stages.
static int panels_rebuild_finish(Panels *P, phf_panels_timing *timing)
{
    if (!P->pending) return PHF_OK;
    P->pending = false;
    Lane *L = &P->lane;
    hipEvent_t *ev = L->ev;
    PHF_HIP(hipEventSynchronize(P->ev_done));
    if (timing) {
        float a = 0, b = 0, c = 0, d = 0, e = 0;
        (void)hipEventElapsedTime(&a, ev[0], ev[1]);
        (void)hipEventElapsedTime(&b, ev[1], ev[2]);
        (void)hipEventElapsedTime(&c, ev[2], ev[3]);
        (void)hipEventElapsedTime(&d, ev[3], ev[4]);
        (void)hipEventElapsedTime(&e, ev[0], ev[5]);
        timing->influence_ms = a;
        timing->normal_ms = b;
        timing->lu_ms = c;
        timing->inverse_ms = d;
        timing->total_ms = e;
    }
    return dense_factor_finish(L, &P->f, *P->info_host);
}

// The rebuild for geometry g, complete when this returns.
static int panels_rebuild(Panels *P, const PanelGeomHost &g, phf_panels_timing *timing)
{
    PHF_TRY(panels_rebuild_finish(P));  // an earlier one still in flight
    PHF_TRY(panels_rebuild_enqueue(P, g));
    return panels_rebuild_finish(P, timing);
}

}  // namespace phf

// ------------------------------------------------------------ C entry ----
using namespace phf;

struct phf_panels_s : public Panels {};

extern "C" {

int phf_panels_create(const double *xs, const double *ys, const double *xe, const double *ye,
                      int64_t m, int32_t problem, double gTotal, phf_panels **out)
{
    clear_error();
    PHF_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    PHF_REQUIRE(m >= 1 && m <= 4096, "panel count out of range (%lld)", (long long)m);
    PHF_REQUIRE(problem == PHF_PANELS_FIXED || problem == PHF_PANELS_PRESCRIBED ||
                    problem == PHF_PANELS_MOVING,
                "unknown problem type %d", problem);
    PHF_REQUIRE(std::isfinite(gTotal), "gTotal must be finite");
    PanelGeomHost g;
    PHF_TRY(panel_geometry_host(xs, ys, xe, ye, m, &g));
    PHF_TRY(Context::get().init());
    PHF_HIP(hipSetDevice(Context::get().device()));
    phf_panels_s *P = new phf_panels_s();
    hipError_t e = hipStreamCreateWithFlags(&P->lane.stream, hipStreamNonBlocking);
    for (int k = 0; e == hipSuccess && k < 2 * PHF_STAGE_COUNT; ++k)
        e = hipEventCreate(&P->lane.ev[k]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&P->ev_geom, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&P->ev_done, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void **>(&P->geom_pin), sizeof(double) * 9 * m);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void **>(&P->info_host), sizeof(int32_t));
    if (e != hipSuccess) {
        set_error("stream/event/pinned buffer creation failed: %s", hipGetErrorName(e));
        phf_panels_destroy(P);
        return PHF_ERR_DEVICE;
    }
    int rc = [&]() -> int {
        Lane *L = &P->lane;
        P->m = m;
        P->problem = problem;
        P->gTotal = gTotal;
        double **bufs[9] = {&P->xs, &P->ys, &P->xe, &P->ye, &P->cosA, &P->sinA, &P->L, &P->xcp,
                            &P->ycp};
        for (double **b : bufs) PHF_ALLOC(*b, L->arena, double, (size_t)m);
        const int64_t rows = problem == PHF_PANELS_PRESCRIBED ? m + 1 : m;
        PHF_ALLOC(P->A, L->arena, double, (size_t)(rows * m));
        PHF_ALLOC(P->s, L->arena, double, (size_t)m);
        PHF_TRY(panels_rebuild(P, g, nullptr));
        zero_doubles<<<div_up(m, 256), 256, 0, L->stream>>>(P->s, m);
        PHF_HIP(hipGetLastError());
        PHF_HIP(hipStreamSynchronize(L->stream));
        return PHF_OK;
    }();
    if (rc != PHF_OK) {
        phf_panels_destroy(P);
        return rc;
    }
    *out = P;
    return PHF_OK;
}

int phf_panels_update(phf_panels *P, const double *xs, const double *ys, const double *xe,
                      const double *ye, phf_panels_timing *timing)
{
    clear_error();
    PHF_REQUIRE(P != nullptr, "panels is NULL");
    PanelGeomHost g;
    PHF_TRY(panel_geometry_host(xs, ys, xe, ye, P->m, &g));
    std::lock_guard<std::mutex> lk(P->mu);
    PHF_HIP(hipSetDevice(Context::get().device()));
    phf_panels_timing t{};
    PHF_TRY(panels_rebuild(P, g, &t));
    // strengths of the old geometry are no start for the new one
    P->have_solution = false;
    if (timing) *timing = t;
    return PHF_OK;
}

int phf_panels_update_async(phf_panels *P, const double *xs, const double *ys, const double *xe,
                            const double *ye)
{
    clear_error();
    PHF_REQUIRE(P != nullptr, "panels is NULL");
    PanelGeomHost g;
    PHF_TRY(panel_geometry_host(xs, ys, xe, ye, P->m, &g));
    std::lock_guard<std::mutex> lk(P->mu);
    PHF_HIP(hipSetDevice(Context::get().device()));
    PHF_TRY(panels_rebuild_finish(P));  // the previous update's errors surface here
    PHF_TRY(panels_rebuild_enqueue(P, g));
    P->have_solution = false;
    return PHF_OK;
}

int phf_panels_destroy(phf_panels *P)
{
    clear_error();
    if (!P) return PHF_OK;
    {
        std::lock_guard<std::mutex> lk(P->mu);
        if (P->lane.stream) (void)hipStreamSynchronize(P->lane.stream);
        fmm_graphs_release(&P->lane);
        P->lane.arena.release();
        P->lane.pin.release();
        for (auto &ev : P->lane.ev)
            if (ev) (void)hipEventDestroy(ev);
        if (P->ev_geom) (void)hipEventDestroy(P->ev_geom);
        if (P->ev_done) (void)hipEventDestroy(P->ev_done);
        if (P->geom_pin) (void)hipHostFree(P->geom_pin);
        if (P->info_host) (void)hipHostFree(P->info_host);
        if (P->lane.stream) (void)hipStreamDestroy(P->lane.stream);
    }
    delete static_cast<phf_panels_s *>(P);
    return PHF_OK;
}

int phf_panels_info(phf_panels *P, int64_t *m, double *xcp, double *ycp, double *A)
{
    clear_error();
    PHF_REQUIRE(P != nullptr, "panels is NULL");
    std::lock_guard<std::mutex> lk(P->mu);
    PHF_TRY(panels_rebuild_finish(P));
    if (m) *m = P->m;
    if (xcp) std::memcpy(xcp, P->hxcp.data(), sizeof(double) * P->m);
    if (ycp) std::memcpy(ycp, P->hycp.data(), sizeof(double) * P->m);
    if (A) {
        PHF_HIP(hipMemcpyAsync(A, P->A, sizeof(double) * P->m * P->m, hipMemcpyDeviceToHost,
                               P->lane.stream));
        PHF_HIP(hipStreamSynchronize(P->lane.stream));
    }
    return PHF_OK;
}

int phf_panels_solve(phf_panels *P, const double *vx, const double *vy, double *strength,
                     int32_t *iterations)
{
    clear_error();
    PHF_REQUIRE(P != nullptr, "panels is NULL");
    PHF_TRY(check_host_array(vx, P->m, "vxExternal"));
    PHF_TRY(check_host_array(vy, P->m, "vyExternal"));
    PHF_REQUIRE(strength != nullptr, "strength is NULL");
    std::lock_guard<std::mutex> lk(P->mu);
    PHF_TRY(panels_rebuild_finish(P));
    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *L = g.lane;
    PHF_TRY(reset_arena(L));
    double *dvx, *dvy;
    PHF_TRY(upload(L, vx, P->m, &dvx));
    PHF_TRY(upload(L, vy, P->m, &dvy));
    PHF_TRY(panels_solve_device(L, P, dvx, dvy, 0.0, 0.0));
    PHF_HIP(hipMemcpyAsync(strength, P->s, sizeof(double) * P->m, hipMemcpyDeviceToHost,
                           L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    if (iterations) *iterations = P->last_iterations;
    return PHF_OK;
}

int phf_system_step_rk2(phf_system *h, const phf_vel_params *vp, phf_panels *P,
                        const phf_step_params *sp, phf_step_info *info)
{
    clear_error();
    PHF_REQUIRE(h != nullptr && sp != nullptr, "NULL argument");
    PHF_TRY(check_vel_params(vp));
    PHF_REQUIRE(std::isfinite(sp->dt) && std::isfinite(sp->vinf_x) && std::isfinite(sp->vinf_y),
                "dt and vInf must be finite");
    if (sp->remesh)
        PHF_REQUIRE(std::isfinite(sp->remesh_sigma) && sp->remesh_sigma > 0 &&
                        std::isfinite(sp->remesh_overlap) && sp->remesh_overlap > 0 &&
                        std::isfinite(sp->remesh_x0) && std::isfinite(sp->remesh_y0) &&
                        std::isfinite(sp->remesh_c) &&
                        (sp->remesh_kernel == PHF_REMESH_M4 ||
                         (sp->remesh_kernel == PHF_REMESH_TUTTY && sp->remesh_c >= 0)),
                    "remesh parameters must be finite (sigma, overlap > 0; a known kernel; "
                    "hNu >= 0 for Tutty)");
    System *s = reinterpret_cast<System *>(h);
    std::lock_guard<std::mutex> lk(s->mu);
    std::unique_lock<std::mutex> lkp;
    if (P) lkp = std::unique_lock<std::mutex>(P->mu);
    PHF_HIP(hipSetDevice(Context::get().device()));
    Lane *L = &s->lane;
    PHF_TRY(reset_arena(L));
    auto t0 = std::chrono::steady_clock::now();
    // stage times from events on the step's stream: no synchronisation
    // between convection and remesh just to time them
    for (auto &ev : L->step_ev)
        if (!ev) PHF_HIP(hipEventCreate(&ev));
    PHF_HIP(hipEventRecord(L->step_ev[0], L->stream));
    StageTimer tm(L, false);
    bool used[PHF_STAGE_COUNT] = {};
    const int64_t n = s->n;
    phf_step_info inf{};
    inf.n_before = n;
    if (n > 0) {
        double *k1x, *k1y, *k2x, *k2y, *x1, *y1;
        PHF_ALLOC(k1x, L->arena, double, (size_t)n);
        PHF_ALLOC(k1y, L->arena, double, (size_t)n);
        PHF_ALLOC(k2x, L->arena, double, (size_t)n);
        PHF_ALLOC(k2y, L->arena, double, (size_t)n);
        PHF_ALLOC(x1, L->arena, double, (size_t)n);
        PHF_ALLOC(y1, L->arena, double, (size_t)n);
        PHF_TRY(stage_velocity(L, s->x, s->y, s->w, n, *vp, P, P != nullptr, sp->vinf_x,
                               sp->vinf_y, k1x, k1y, tm, used));
        euler_stage<<<div_up(n, 256), 256, 0, L->stream>>>(s->x, s->y, k1x, k1y, n, sp->dt, x1,
                                                          y1);
        PHF_TRY(stage_velocity(L, x1, y1, s->w, n, *vp, P, P != nullptr && sp->panel_update == 1,
                               sp->vinf_x, sp->vinf_y, k2x, k2y, tm, used));
        heun_final<<<div_up(n, 256), 256, 0, L->stream>>>(s->x, s->y, k1x, k1y, k2x, k2y, n,
                                                         0.5 * sp->dt);
        PHF_HIP(hipGetLastError());
        if (P) inf.solve_iterations = P->last_iterations;
    }
    PHF_HIP(hipEventRecord(L->step_ev[1], L->stream));
    double4 *dm;
    PHF_ALLOC(dm, L->arena, double4, 2);
    double4 hm[2] = {};
    if (sp->remesh && n > 0) {
        PHF_TRY(moments_device(L, s->x, s->y, s->w, n, dm));
        RemeshOut ro;  // the moments come back with the remesh's output count
        ro.read_src = dm;
        ro.read_bytes = sizeof(double4);
        ro.read_dst = &hm[0];
        PHF_TRY(remesh_device(L, s->x, s->y, s->w, n, sp->remesh_sigma, sp->remesh_overlap,
                              sp->remesh_x0, sp->remesh_y0, sp->remesh_c, &ro,
                              sp->remesh_kernel));
        PHF_TRY(sys_reserve(s, ro.m));
        if (ro.m > 0) {
            copy3_f64<<<(unsigned)std::min<int64_t>(div_up(ro.m, 256), 4096), 256, 0, L->stream>>>(
                ro.x, ro.y, ro.w, ro.m, s->x, s->y, s->w);
            PHF_HIP(hipGetLastError());
        }
        s->n = ro.m;
        if (sp->pc_enabled) {
            // Blobs.evolve: redistribute, then populationControl (__blobsClass.py:771-781)
            double *px, *py, *pw;
            PHF_ALLOC(px, L->arena, double, (size_t)std::max<int64_t>(s->n, 1));
            PHF_ALLOC(py, L->arena, double, (size_t)std::max<int64_t>(s->n, 1));
            PHF_ALLOC(pw, L->arena, double, (size_t)std::max<int64_t>(s->n, 1));
            int64_t m = 0;
            PHF_TRY(population_control_device(L, s->x, s->y, s->w, s->n, sp->pc_w_threshold,
                                              sp->pc_total_threshold, px, py, pw, &m, nullptr));
            PHF_HIP(hipMemcpyAsync(s->x, px, sizeof(double) * m, hipMemcpyDeviceToDevice, L->stream));
            PHF_HIP(hipMemcpyAsync(s->y, py, sizeof(double) * m, hipMemcpyDeviceToDevice, L->stream));
            PHF_HIP(hipMemcpyAsync(s->w, pw, sizeof(double) * m, hipMemcpyDeviceToDevice, L->stream));
            s->n = m;
        }
        PHF_TRY(moments_device(L, s->x, s->y, s->w, s->n, dm + 1));
        PHF_HIP(hipMemcpyAsync(&hm[1], dm + 1, sizeof(double4), hipMemcpyDeviceToHost, L->stream));
    } else if (n > 0) {
        PHF_TRY(moments_device(L, s->x, s->y, s->w, n, dm));
        PHF_HIP(hipMemcpyAsync(&hm[0], dm, sizeof(double4), hipMemcpyDeviceToHost, L->stream));
        hm[1] = hm[0];
    }
    PHF_HIP(hipEventRecord(L->step_ev[2], L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    s->has_velocity = false;
    auto t2 = std::chrono::steady_clock::now();
    inf.n_after = s->n;
    moments_to(hm[0], &inf.circ_before, &inf.impulse_x_before, &inf.impulse_y_before,
               &inf.m2_before);
    moments_to(hm[1], &inf.circ_after, &inf.impulse_x_after, &inf.impulse_y_after,
               &inf.m2_after);
    float ms_conv = 0.f, ms_rem = 0.f;
    PHF_HIP(hipEventElapsedTime(&ms_conv, L->step_ev[0], L->step_ev[1]));
    PHF_HIP(hipEventElapsedTime(&ms_rem, L->step_ev[1], L->step_ev[2]));
    inf.ms_convection = ms_conv;
    inf.ms_remesh = ms_rem;
    inf.ms_total = std::chrono::duration<double, std::milli>(t2 - t0).count();
    if (info) *info = inf;
    return PHF_OK;
}

}  // extern "C"
