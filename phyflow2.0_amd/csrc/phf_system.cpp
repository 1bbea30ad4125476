// phf_system.cpp -- device-resident blob systems (one stream + workspace
// each), velocity evaluation on resident data.
#include "phf_system.hpp"

#include <cmath>
#include <mutex>

namespace phf {

// Capacity for n blobs (contents not kept).  A system that outgrows its
// buffers (a remesh that adds blobs) gets half again as many, so that a
// slowly growing simulation does not reallocate every step.
int sys_reserve(System *s, int64_t n)
{
    if (n <= s->cap) return PHF_OK;
    // work enqueued on the system's streams may still read the old buffers
    PHF_TRY(lane_quiesce(&s->lane));
    double *bufs[5] = {s->x, s->y, s->w, s->vx, s->vy};
    const bool grow = s->cap > 0;
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    s->x = s->y = s->w = s->vx = s->vy = nullptr;
    const int64_t c = std::max<int64_t>(grow ? std::max(n, s->cap + s->cap / 2) : n, 1);
    s->cap = 0;
    PHF_HIP(hipMalloc(&s->x, sizeof(double) * c));
    PHF_HIP(hipMalloc(&s->y, sizeof(double) * c));
    PHF_HIP(hipMalloc(&s->w, sizeof(double) * c));
    PHF_HIP(hipMalloc(&s->vx, sizeof(double) * c));
    PHF_HIP(hipMalloc(&s->vy, sizeof(double) * c));
    s->cap = c;
    return PHF_OK;
}

static int sys_upload(System *s, const double *x, const double *y, const double *w, int64_t n)
{
    PHF_TRY(sys_reserve(s, n));
    s->n = n;
    s->has_velocity = false;
    if (n > 0) {
        PHF_HIP(hipMemcpyAsync(s->x, x, sizeof(double) * n, hipMemcpyHostToDevice, s->lane.stream));
        PHF_HIP(hipMemcpyAsync(s->y, y, sizeof(double) * n, hipMemcpyHostToDevice, s->lane.stream));
        PHF_HIP(hipMemcpyAsync(s->w, w, sizeof(double) * n, hipMemcpyHostToDevice, s->lane.stream));
    }
    PHF_HIP(hipStreamSynchronize(s->lane.stream));
    return PHF_OK;
}

int check_vel_params(const phf_vel_params *p)
{
    PHF_REQUIRE(p != nullptr, "params are NULL");
    PHF_REQUIRE(p->method == PHF_METHOD_FMM || p->method == PHF_METHOD_DIRECT,
                "unknown method %d", p->method);
    PHF_REQUIRE(p->prec == PHF_PREC_F64 || p->prec == PHF_PREC_F32C, "unknown precision %d",
                p->prec);
    PHF_REQUIRE(std::isfinite(p->sigma) && p->sigma > 0, "sigma must be finite and > 0 (got %g)",
                p->sigma);
    PHF_REQUIRE(std::isfinite(p->k) && p->k > 0, "k must be finite and > 0 (got %g)", p->k);
    PHF_REQUIRE(std::isfinite(p->tol) && p->tol >= 0, "tol must be finite and >= 0 (got %g)",
                p->tol);
    PHF_REQUIRE(std::isfinite(p->cutoff), "cutoff must be finite");
    PHF_REQUIRE(std::isfinite(p->leaf_side) && p->leaf_side <= 16.0,
                "leaf_side must be finite and <= 16 (got %g)", p->leaf_side);
    PHF_REQUIRE(p->order <= 40, "order must be <= 40 (got %d)", p->order);
    PHF_REQUIRE(p->ws <= 3, "ws must be <= 3 (got %d)", p->ws);
    return PHF_OK;
}

// Velocity induced by blobs (x, y, w)[n] at targets (xt, yt)[nt] in the
// reference's velocity() convention, into (vx, vy) (device).  same: the
// targets are the blobs themselves.
int velocity_device(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                    const double *xt, const double *yt, int64_t nt, bool same,
                    const phf_vel_params &p, double *vx, double *vy, StageTimer &tm, bool *used,
                    FmmInfo *fi)
{
    if (nt == 0) return PHF_OK;
    const double ks = p.k * p.sigma * p.sigma;
    if (p.method == PHF_METHOD_DIRECT) {
        tm.begin(PHF_STAGE_P2P);
        used[PHF_STAGE_P2P] = true;
        PairParams pp = make_gauss_params(ks);
        PHF_TRY(direct_device(L, PAIR_GAUSS_EPS, p.prec, pp, x, y, w, n, xt, yt, nt, vx, vy));
        tm.end(PHF_STAGE_P2P);
        return PHF_OK;
    }
    const double xopt = std::sqrt(1.2564312086261696770 * ks);  // _induced.py:236-239
    const double cutoff = p.cutoff > 0 ? p.cutoff : 5.0 * xopt;
    double *qr, *qi;
    PHF_ALLOC(qr, L->arena, double, (size_t)nt);
    PHF_ALLOC(qi, L->arena, double, (size_t)nt);
    if (p.tol == 0.0) {
        tm.begin(PHF_STAGE_P2P);
        used[PHF_STAGE_P2P] = true;
        PairParams pp = make_oseen_params(xopt, cutoff);
        PHF_TRY(direct_device(L, PAIR_OSEEN, p.prec, pp, x, y, w, n, xt, yt, nt, qr, qi));
        tm.end(PHF_STAGE_P2P);
    } else {
        FmmArgs a{};
        a.xb = x; a.yb = y; a.wb = w; a.nb = n;
        a.xt = xt; a.yt = yt; a.nt = nt; a.same = same;
        a.xopt = xopt; a.cutoff = cutoff;
        a.p = p.order > 0 ? p.order : fmm_order_from_tol(p.tol);
        a.ws = p.ws > 0 ? p.ws : 1;
        a.prec = p.prec;
        a.Lhint = fmm_level(n, p.leaf, p.ndirect > 0 ? p.ndirect : 35);
        a.leaf_side = p.leaf_side;
        a.qr = vx;
        a.qi = vy;
        a.velocity_out = true;
        return fmm_device(L, a, tm, used, fi);
    }
    return q_to_velocity_device(L, qr, qi, nt, vx, vy);
}

int system_velocity_device(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                           const phf_vel_params &p, double *vx, double *vy, StageTimer &tm,
                           bool *used, FmmInfo *fi)
{
    return velocity_device(L, x, y, w, n, x, y, n, true, p, vx, vy, tm, used, fi);
}

}  // namespace phf

using namespace phf;

extern "C" {

int phf_system_create(const double *x, const double *y, const double *w, int64_t n,
                      phf_system **out)
{
    clear_error();
    PHF_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    PHF_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "n out of range (%lld)", (long long)n);
    PHF_TRY(check_host_array(x, n, "x"));
    PHF_TRY(check_host_array(y, n, "y"));
    PHF_TRY(check_host_array(w, n, "w"));
    PHF_TRY(Context::get().init());
    PHF_HIP(hipSetDevice(Context::get().device()));
    System *s = new System();
    hipError_t e = hipStreamCreateWithFlags(&s->lane.stream, hipStreamNonBlocking);
    for (int k = 0; e == hipSuccess && k < 2 * PHF_STAGE_COUNT; ++k)
        e = hipEventCreate(&s->lane.ev[k]);
    if (e != hipSuccess) {
        set_error("stream/event creation failed: %s", hipGetErrorName(e));
        delete s;
        return PHF_ERR_DEVICE;
    }
    int rc = sys_upload(s, x, y, w, n);
    if (rc != PHF_OK) {
        phf_system_destroy(reinterpret_cast<phf_system *>(s));
        return rc;
    }
    *out = reinterpret_cast<phf_system *>(s);
    return PHF_OK;
}

int phf_system_destroy(phf_system *h)
{
    clear_error();
    if (!h) return PHF_OK;
    System *s = reinterpret_cast<System *>(h);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->lane.stream) (void)hipStreamSynchronize(s->lane.stream);
        double *bufs[5] = {s->x, s->y, s->w, s->vx, s->vy};
        for (double *b : bufs)
            if (b) (void)hipFree(b);
        fmm_graphs_release(&s->lane);
        s->lane.arena.release();
        s->lane.pin.release();
        for (auto &ev : s->lane.ev)
            if (ev) (void)hipEventDestroy(ev);
        for (auto &ev : s->lane.step_ev)
            if (ev) (void)hipEventDestroy(ev);
        if (s->lane.stream) (void)hipStreamDestroy(s->lane.stream);
    }
    delete s;
    return PHF_OK;
}

int phf_system_size(phf_system *h, int64_t *n)
{
    clear_error();
    PHF_REQUIRE(h && n, "NULL argument");
    *n = reinterpret_cast<System *>(h)->n;
    return PHF_OK;
}

int phf_system_get(phf_system *h, double *x, double *y, double *w)
{
    clear_error();
    PHF_REQUIRE(h != nullptr, "system is NULL");
    System *s = reinterpret_cast<System *>(h);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->n > 0) {
        if (x) PHF_HIP(hipMemcpyAsync(x, s->x, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->lane.stream));
        if (y) PHF_HIP(hipMemcpyAsync(y, s->y, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->lane.stream));
        if (w) PHF_HIP(hipMemcpyAsync(w, s->w, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->lane.stream));
    }
    PHF_HIP(hipStreamSynchronize(s->lane.stream));
    return PHF_OK;
}

int phf_system_set(phf_system *h, const double *x, const double *y, const double *w, int64_t n)
{
    clear_error();
    PHF_REQUIRE(h != nullptr, "system is NULL");
    PHF_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "n out of range (%lld)", (long long)n);
    PHF_TRY(check_host_array(x, n, "x"));
    PHF_TRY(check_host_array(y, n, "y"));
    PHF_TRY(check_host_array(w, n, "w"));
    System *s = reinterpret_cast<System *>(h);
    std::lock_guard<std::mutex> lk(s->mu);
    PHF_HIP(hipSetDevice(Context::get().device()));
    return sys_upload(s, x, y, w, n);
}

int phf_system_velocity(phf_system *h, const phf_vel_params *p, double *vx, double *vy,
                        phf_eval_info *info)
{
    clear_error();
    PHF_REQUIRE(h != nullptr, "system is NULL");
    PHF_TRY(check_vel_params(p));
    System *s = reinterpret_cast<System *>(h);
    std::lock_guard<std::mutex> lk(s->mu);
    PHF_HIP(hipSetDevice(Context::get().device()));
    Lane *L = &s->lane;
    const bool timed = info != nullptr || timing_enabled();
    StageTimer tm(L, timed);
    if (info != nullptr && (info->flags & PHF_INFO_TIME_P2P))
        tm.mask = 1u << PHF_STAGE_P2P;
    bool used[PHF_STAGE_COUNT] = {};
    FmmInfo fi;
    unsigned long long counts[2] = {0, 0};
    // resident buffers keep their addresses: the FMM launch sequence is
    // recorded once and replayed; a replay whose points needed another
    // domain is made again after a fresh recording
    for (int attempt = 0; attempt < 2; ++attempt) {
        PHF_TRY(reset_arena(L));
        for (bool &u : used) u = false;
        used[PHF_STAGE_TOTAL] = true;
        fi = FmmInfo{};
        fi.want_counts = info != nullptr && (info->flags & PHF_INFO_COUNT_WORK) &&
                         p->method == PHF_METHOD_FMM && p->tol > 0;
        fi.graph = !(info != nullptr && (info->flags & PHF_INFO_EAGER));
        fi.no_speculation = attempt > 0;
        tm.begin(PHF_STAGE_TOTAL);
        PHF_TRY(system_velocity_device(L, s->x, s->y, s->w, s->n, *p, s->vx, s->vy, tm, used,
                                       &fi));
        if (vx && vy && s->n > 0) {
            tm.begin(PHF_STAGE_DOWNLOAD);
            used[PHF_STAGE_DOWNLOAD] = true;
            PHF_HIP(hipMemcpyAsync(vx, s->vx, sizeof(double) * s->n, hipMemcpyDeviceToHost,
                                   L->stream));
            PHF_HIP(hipMemcpyAsync(vy, s->vy, sizeof(double) * s->n, hipMemcpyDeviceToHost,
                                   L->stream));
            tm.end(PHF_STAGE_DOWNLOAD);
        }
        tm.end(PHF_STAGE_TOTAL);
        if (fi.want_counts && fi.d_counts)
            PHF_HIP(hipMemcpyAsync(counts, fi.d_counts, sizeof(counts), hipMemcpyDeviceToHost,
                                   L->stream));
        PHF_HIP(hipStreamSynchronize(L->stream));
        if (!fi.domain_ok || fi.domain_ok()) break;
    }
    s->has_velocity = true;
    for (int k = 0; k < PHF_STAGE_COUNT; ++k) used[k] = used[k] && tm.timed(k);
    record_stage_times(L, used);
    if (info) {
        info->L = fi.L;
        info->p = fi.p;
        info->nnear = fi.nnear;
        info->D = fi.D;
        info->p2p_pairs = p->method == PHF_METHOD_FMM && p->tol > 0
                              ? (fi.want_counts ? (int64_t)counts[0] : -1)
                              : s->n * s->n;
        info->m2l_links = (int64_t)counts[1];
        for (int k = 0; k < PHF_STAGE_COUNT; ++k) {
            float ms = -1.0f;
            if (used[k] && hipEventElapsedTime(&ms, L->ev[2 * k], L->ev[2 * k + 1]) == hipSuccess)
                info->ms[k] = ms;
            else
                info->ms[k] = -1.0;
        }
    }
    return PHF_OK;
}

int phf_system_get_velocity(phf_system *h, double *vx, double *vy)
{
    clear_error();
    PHF_REQUIRE(h && vx && vy, "NULL argument");
    System *s = reinterpret_cast<System *>(h);
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->has_velocity) {
        set_error("no velocity has been evaluated for this system");
        return PHF_ERR_INVALID;
    }
    if (s->n > 0) {
        PHF_HIP(hipMemcpyAsync(vx, s->vx, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->lane.stream));
        PHF_HIP(hipMemcpyAsync(vy, s->vy, sizeof(double) * s->n, hipMemcpyDeviceToHost, s->lane.stream));
    }
    PHF_HIP(hipStreamSynchronize(s->lane.stream));
    return PHF_OK;
}

}  // extern "C"
