// phf_common.cpp -- error state, device context, lane pool, arenas, timers.
#include "phf_common.hpp"

#include <cmath>
#include <cstdarg>

namespace phf {

static thread_local char g_err[1024] = "";
static thread_local double g_stage_ms[PHF_STAGE_COUNT];
static thread_local bool g_stage_init = false;
static bool g_timing = false;

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

// ----------------------------------------------------------------- arena --
static size_t round_mb(size_t b)
{
    const size_t mb = size_t(1) << 20;
    return (b + mb - 1) / mb * mb;
}

int Arena::reset()
{
    size_t pass = used;
    if (pass > peak) peak = pass;
    used = 0;
    cur = 0;
    top = 0;
    if (blocks.size() > 1) {
        size_t total = 0;
        for (auto &b : blocks) total += b.cap;
        release();
        char *p = nullptr;
        PHF_HIP(hipMalloc(&p, total));
        blocks.push_back({p, total});
    }
    return PHF_OK;
}

void *Arena::alloc_bytes(size_t bytes)
{
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    while (cur < blocks.size()) {
        if (top + bytes <= blocks[cur].cap) {
            void *p = blocks[cur].base + top;
            top += bytes;
            used += bytes;
            return p;
        }
        ++cur;
        top = 0;
    }
    size_t last = blocks.empty() ? 0 : blocks.back().cap;
    size_t want = round_mb(bytes > 2 * last ? bytes : 2 * last);
    if (want < (size_t(64) << 20)) want = size_t(64) << 20;
    char *p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("device allocation of %zu bytes failed (%s)", want,
                  hipGetErrorName(e));
        return nullptr;
    }
    blocks.push_back({p, want});
    cur = blocks.size() - 1;
    top = bytes;
    used += bytes;
    return p;
}

Arena::Mark Arena::mark() const
{
    Mark m;
    m.nblocks = blocks.size();
    m.cur = cur;
    m.top = top;
    m.used = used;
    m.base0 = blocks.empty() ? nullptr : blocks[0].base;
    return m;
}

bool Arena::at(const Mark &m) const
{
    return blocks.size() == m.nblocks && cur == m.cur && top == m.top &&
           (blocks.empty() ? nullptr : blocks[0].base) == m.base0;
}

void Arena::jump(const Mark &from, const Mark &to)
{
    cur = to.cur;
    top = to.top;
    used += to.used - from.used;
}

void Arena::release()
{
    for (auto &b : blocks) (void)hipFree(b.base);
    blocks.clear();
    cur = 0;
    top = 0;
}

size_t Arena::capacity() const
{
    size_t t = 0;
    for (auto &b : blocks) t += b.cap;
    return t;
}

int Pinned::ensure(size_t bytes)
{
    if (pending) {
        PHF_HIP(hipEventSynchronize(last));
        pending = false;
    }
    if (bytes <= cap) return PHF_OK;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t want = (bytes + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
    PHF_HIP(hipHostMalloc(&ptr, want, hipHostMallocDefault));
    cap = want;
    return PHF_OK;
}

int Pinned::mark(hipStream_t s)
{
    if (!last) PHF_HIP(hipEventCreateWithFlags(&last, hipEventDisableTiming));
    PHF_HIP(hipEventRecord(last, s));
    pending = true;
    return PHF_OK;
}

void Pinned::release()
{
    if (pending) (void)hipEventSynchronize(last);
    pending = false;
    if (ptr) (void)hipHostFree(ptr);
    if (last) (void)hipEventDestroy(last);
    ptr = nullptr;
    last = nullptr;
    cap = 0;
}

int lane_quiesce(Lane *L)
{
    if (L->stream) PHF_HIP(hipStreamSynchronize(L->stream));
    if (L->stream2) PHF_HIP(hipStreamSynchronize(L->stream2));
    return PHF_OK;
}

int reset_arena(Lane *L)
{
    if (L->arena.reset_frees()) PHF_TRY(lane_quiesce(L));
    return L->arena.reset();
}

// --------------------------------------------------------------- context --
static const int kLanes = 4;

Context &Context::get()
{
    static Context *ctx = new Context();  // never destroyed: lanes live until exit
    return *ctx;
}

int Context::init()
{
    std::lock_guard<std::mutex> lk(mu_);
    if (ready_) return PHF_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        set_error("no HIP device available (%s)", hipGetErrorName(e));
        return PHF_ERR_DEVICE;
    }
    PHF_HIP(hipSetDevice(0));
    PHF_HIP(hipGetDeviceProperties(&props_, 0));
    dev_ = 0;
    // device address of the allocation-free direct sum's scratch array
    PHF_TRY(direct_scratch_setup(dev_));
    for (int i = 0; i < kLanes; ++i) {
        Lane *l = new Lane();
        l->index = i;
        PHF_HIP(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking));
        for (int k = 0; k < 2 * PHF_STAGE_COUNT; ++k)
            PHF_HIP(hipEventCreate(&l->ev[k]));
        lanes_.push_back(l);
    }
    ready_ = true;
    return PHF_OK;
}

int Context::acquire(Lane **out)
{
    *out = nullptr;
    PHF_TRY(init());
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        for (Lane *l : lanes_)
            if (!l->busy) {
                l->busy = true;
                *out = l;
                lk.unlock();
                // Each host thread uses the device this library selected.
                PHF_HIP(hipSetDevice(dev_));
                return PHF_OK;
            }
        cv_.wait(lk);
    }
}

void Context::release(Lane *l)
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        l->busy = false;
    }
    cv_.notify_one();
}

// ---------------------------------------------------------------- timing --
void reset_stage_times()
{
    for (int i = 0; i < PHF_STAGE_COUNT; ++i) g_stage_ms[i] = -1.0;
    g_stage_init = true;
}

void record_stage_times(Lane *lane, const bool *used)
{
    reset_stage_times();
    if (!g_timing) return;
    for (int s = 0; s < PHF_STAGE_COUNT; ++s) {
        if (!used[s]) continue;
        float ms = -1.0f;
        if (hipEventElapsedTime(&ms, lane->ev[2 * s], lane->ev[2 * s + 1]) ==
            hipSuccess)
            g_stage_ms[s] = ms;
    }
}

void get_stage_times(double *ms_out)
{
    if (!g_stage_init) reset_stage_times();
    for (int i = 0; i < PHF_STAGE_COUNT; ++i) ms_out[i] = g_stage_ms[i];
}

bool timing_enabled() { return g_timing; }
void set_timing(bool on) { g_timing = on; }

// ------------------------------------------------------------ validation --
int check_host_array(const double *p, int64_t n, const char *name, bool finite)
{
    if (n > 0 && p == nullptr) {
        set_error("%s is NULL but has %lld entries", name, (long long)n);
        return PHF_ERR_INVALID;
    }
    if (finite)
        for (int64_t i = 0; i < n; ++i)
            if (!std::isfinite(p[i])) {
                set_error("%s[%lld] is not finite", name, (long long)i);
                return PHF_ERR_INVALID;
            }
    return PHF_OK;
}

}  // namespace phf

// ------------------------------------------------------------- C entry ---
extern "C" {

const char *phf_last_error(void) { return phf::g_err; }

int phf_init(void)
{
    phf::clear_error();
    return phf::Context::get().init();
}

int phf_device_info(char *name, int64_t name_len, int32_t *num_cu,
                    int64_t *hbm_bytes)
{
    phf::clear_error();
    PHF_TRY(phf::Context::get().init());
    const hipDeviceProp_t &p = phf::Context::get().props();
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (num_cu) *num_cu = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
    return PHF_OK;
}

int phf_synchronize(void)
{
    phf::clear_error();
    PHF_TRY(phf::Context::get().init());
    PHF_HIP(hipSetDevice(phf::Context::get().device()));
    PHF_HIP(hipDeviceSynchronize());
    return PHF_OK;
}

int phf_set_timing(int32_t enabled)
{
    phf::set_timing(enabled != 0);
    return PHF_OK;
}

int phf_last_stage_times(double *ms)
{
    if (!ms) {
        phf::set_error("ms is NULL");
        return PHF_ERR_INVALID;
    }
    phf::get_stage_times(ms);
    return PHF_OK;
}

}  // extern "C"
