// phf_serve.cpp -- request loop over device-resident simulations.
//
// There is no serving layer in the reference (each Python caller drives its
// own Blobs object synchronously).  This server keeps many simulations
// resident in HBM and advances them on request:
//   * admission: a simulation is registered only if its HBM footprint fits
//     the server's budget (PHF_ERR_BUSY otherwise); the footprint grows to
//     the measured workspace after each step;
//   * continuous batching: requests are queued on arrival and handed to
//     `slots` worker threads as soon as a slot frees; every simulation owns
//     a HIP stream, so the steps of different simulations overlap on the
//     device while the queue keeps filling;
//   * requests for one simulation run in submission order (never two at
//     once); a step's result does not depend on what else is running;
//   * per-request latency = completion time - submission time.
#include "phf_system.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>
#include <unordered_map>
#include <vector>

namespace phf {

using Clock = std::chrono::steady_clock;

struct ServeRequest {
    int64_t id = 0, sim = 0;
    phf_vel_params vp{};
    phf_step_params sp{};
    Clock::time_point t_submit, t_start, t_end;
    int status = 1;  // 1 queued, 2 running, 0 done, <0 error
    phf_step_info info{};
    std::string error;
};

struct ServeSim {
    phf_system *sys = nullptr;
    int64_t footprint = 0;
    bool busy = false;
    std::deque<int64_t> pending;  // request ids in submission order
};

struct Server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    int slots = 1;
    int64_t budget = 0, used = 0;
    int64_t next_sim = 1, next_req = 1;
    std::map<int64_t, ServeSim> sims;
    std::unordered_map<int64_t, ServeRequest> reqs;
    std::deque<int64_t> ready;  // simulations with pending work and not busy
    std::vector<std::thread> workers;
    // statistics
    std::vector<double> lat_ms;
    int64_t completed = 0, failed = 0;
    Clock::time_point t_first, t_last;
    bool have_first = false;
};

int64_t footprint_estimate(int64_t n)
{
    // resident arrays (x, y, w, vx, vy) + measured FMM/remesh workspace per
    // blob on the F32C path (~250 B) with margin
    return n * (5 * 8 + 320) + (int64_t(64) << 20);
}

static void worker_loop(Server *S)
{
    for (;;) {
        int64_t sim_id = 0, rid = 0;
        ServeRequest req;
        phf_system *sys = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] { return S->stop || !S->ready.empty(); });
            if (S->stop && S->ready.empty()) return;
            sim_id = S->ready.front();
            S->ready.pop_front();
            ServeSim &sm = S->sims[sim_id];
            sm.busy = true;
            rid = sm.pending.front();
            sm.pending.pop_front();
            ServeRequest &r = S->reqs[rid];
            r.status = 2;
            r.t_start = Clock::now();
            req = r;
            sys = sm.sys;
        }
        phf_step_info info{};
        int rc = phf_system_step_rk2(sys, &req.vp, nullptr, &req.sp, &info);
        std::string err = rc == PHF_OK ? "" : phf_last_error();
        int64_t fp = 0;
        if (rc == PHF_OK) {
            const System *s = reinterpret_cast<const System *>(sys);
            fp = 5 * 8 * s->cap + (int64_t)s->lane.arena.capacity();
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            ServeRequest &r = S->reqs[rid];
            r.t_end = Clock::now();
            r.status = rc == PHF_OK ? 0 : rc;
            r.info = info;
            r.error = err;
            ServeSim &sm = S->sims[sim_id];
            if (fp > sm.footprint) {
                S->used += fp - sm.footprint;
                sm.footprint = fp;
            }
            sm.busy = false;
            if (!sm.pending.empty()) S->ready.push_back(sim_id);
            if (rc == PHF_OK) {
                ++S->completed;
                S->lat_ms.push_back(
                    std::chrono::duration<double, std::milli>(r.t_end - r.t_submit).count());
            } else {
                ++S->failed;
            }
            if (!S->have_first || r.t_submit < S->t_first) S->t_first = r.t_submit;
            S->have_first = true;
            if (r.t_end > S->t_last) S->t_last = r.t_end;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_one();
    }
}

}  // namespace phf

using namespace phf;

struct phf_server_s : public Server {};

extern "C" {

int phf_server_create(int32_t slots, int64_t hbm_budget_bytes, phf_server **out)
{
    clear_error();
    PHF_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    PHF_REQUIRE(slots >= 1 && slots <= 64, "slots must be in [1, 64] (got %d)", slots);
    PHF_TRY(Context::get().init());
    int64_t total = (int64_t)Context::get().props().totalGlobalMem;
    int64_t budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : total - total / 8;
    phf_server *S = new phf_server();
    S->slots = slots;
    S->budget = budget;
    for (int i = 0; i < slots; ++i) S->workers.emplace_back(worker_loop, S);
    *out = S;
    return PHF_OK;
}

int phf_server_destroy(phf_server *S)
{
    clear_error();
    if (!S) return PHF_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers) t.join();
    delete S;  // systems stay owned by the caller
    return PHF_OK;
}

int phf_server_add_system(phf_server *S, phf_system *sys, int64_t *id)
{
    clear_error();
    PHF_REQUIRE(S && sys && id, "NULL argument");
    int64_t n = 0;
    PHF_TRY(phf_system_size(sys, &n));
    const int64_t fp = footprint_estimate(n);
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->used + fp > S->budget) {
        set_error("admission refused: %lld bytes needed, %lld of %lld in use",
                  (long long)fp, (long long)S->used, (long long)S->budget);
        return PHF_ERR_BUSY;
    }
    const int64_t k = S->next_sim++;
    S->sims[k].sys = sys;
    S->sims[k].footprint = fp;
    S->used += fp;
    *id = k;
    return PHF_OK;
}

int phf_server_remove_system(phf_server *S, int64_t id)
{
    clear_error();
    PHF_REQUIRE(S != nullptr, "server is NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->sims.find(id);
    if (it == S->sims.end()) {
        set_error("unknown simulation %lld", (long long)id);
        return PHF_ERR_NOTFOUND;
    }
    S->cv_done.wait(lk, [&] { return !it->second.busy && it->second.pending.empty(); });
    S->used -= it->second.footprint;
    S->sims.erase(it);
    return PHF_OK;
}

int phf_server_submit(phf_server *S, int64_t sim, const phf_vel_params *vp,
                      const phf_step_params *sp, int64_t *req_id)
{
    clear_error();
    PHF_REQUIRE(S && vp && sp && req_id, "NULL argument");
    PHF_TRY(check_vel_params(vp));
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->sims.find(sim);
    if (it == S->sims.end()) {
        set_error("unknown simulation %lld", (long long)sim);
        return PHF_ERR_NOTFOUND;
    }
    const int64_t rid = S->next_req++;
    ServeRequest &r = S->reqs[rid];
    r.id = rid;
    r.sim = sim;
    r.vp = *vp;
    r.sp = *sp;
    r.t_submit = Clock::now();
    const bool was_idle = !it->second.busy && it->second.pending.empty();
    it->second.pending.push_back(rid);
    if (was_idle) S->ready.push_back(sim);
    *req_id = rid;
    S->cv_work.notify_one();
    return PHF_OK;
}

int phf_server_wait(phf_server *S, int64_t req_id, double timeout_s, phf_request_result *res)
{
    clear_error();
    PHF_REQUIRE(S != nullptr, "server is NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(req_id);
    if (it == S->reqs.end()) {
        set_error("unknown request %lld", (long long)req_id);
        return PHF_ERR_NOTFOUND;
    }
    auto pred = [&] { return it->second.status <= 0; };
    if (timeout_s < 0)
        S->cv_done.wait(lk, pred);
    else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
        set_error("request %lld not finished within %g s", (long long)req_id, timeout_s);
        return PHF_ERR_BUSY;
    }
    const ServeRequest &r = it->second;
    if (res) {
        res->status = r.status;
        res->latency_ms = std::chrono::duration<double, std::milli>(r.t_end - r.t_submit).count();
        res->queue_ms = std::chrono::duration<double, std::milli>(r.t_start - r.t_submit).count();
        res->info = r.info;
    }
    if (r.status < 0) {
        set_error("request %lld failed: %s", (long long)req_id, r.error.c_str());
        const int st = r.status;
        S->reqs.erase(it);
        return st;
    }
    S->reqs.erase(it);
    return PHF_OK;
}

int phf_server_drain(phf_server *S)
{
    clear_error();
    PHF_REQUIRE(S != nullptr, "server is NULL");
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] {
        for (auto &kv : S->reqs)
            if (kv.second.status > 0) return false;
        return true;
    });
    return PHF_OK;
}

int phf_server_stats(phf_server *S, int32_t reset, phf_serve_stats *st)
{
    clear_error();
    PHF_REQUIRE(S && st, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<double> v = S->lat_ms;
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) -> double {
        if (v.empty()) return 0.0;
        const double pos = q * (double)(v.size() - 1);
        const size_t i = (size_t)pos;
        const double f = pos - (double)i;
        return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
    };
    st->completed = S->completed;
    st->failed = S->failed;
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = v.empty() ? 0.0 : v.back();
    st->span_s = S->have_first
                     ? std::chrono::duration<double>(S->t_last - S->t_first).count()
                     : 0.0;
    st->hbm_used = S->used;
    st->hbm_budget = S->budget;
    st->slots = S->slots;
    if (reset) {
        S->lat_ms.clear();
        S->completed = S->failed = 0;
        S->have_first = false;
    }
    return PHF_OK;
}

}  // extern "C"
