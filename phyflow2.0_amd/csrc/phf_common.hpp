// phf_common.hpp -- runtime plumbing shared by every translation unit of
// libphyflow_amd: error state, HIP call checking, the per-process device
// context (streams + reusable device workspaces), and small device helpers.
//
// Design notes
//  * Every exported call returns a status code and leaves a message in a
//    thread-local buffer (phf_last_error).  The reference prints CUDA errors
//    and carries on (cudaeval.cu:48-93); here every HIP error is returned.
//  * The reference library is not reentrant (global textures, static
//    workspaces: cudaeval.cu:115-128, fmmshift.cpp:27-31).  Here a call
//    borrows one "lane" (HIP stream + growable device arena + pinned staging)
//    from a small pool, so concurrent host threads run on separate streams.
//  * Arenas are grown, never shrunk, so steady-state calls do no hipMalloc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <condition_variable>
#include <string>
#include <vector>

#include "phyflow_amd.h"

namespace phf {

// ---------------------------------------------------------------- errors --
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

struct Status {
    int code = PHF_OK;
    bool ok() const { return code == PHF_OK; }
};

#define PHF_HIP(call)                                                        \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            ::phf::set_error("HIP error %s (%d) at %s:%d: %s",              \
                             hipGetErrorName(e_), (int)e_, __FILE__,        \
                             __LINE__, #call);                              \
            return (e_ == hipErrorOutOfMemory) ? PHF_ERR_NOMEM              \
                                               : PHF_ERR_DEVICE;            \
        }                                                                    \
    } while (0)

#define PHF_TRY(expr)                                                        \
    do {                                                                     \
        int rc_ = (expr);                                                    \
        if (rc_ != PHF_OK) return rc_;                                       \
    } while (0)

#define PHF_REQUIRE(cond, ...)                                               \
    do {                                                                     \
        if (!(cond)) {                                                       \
            ::phf::set_error(__VA_ARGS__);                                   \
            return PHF_ERR_INVALID;                                          \
        }                                                                    \
    } while (0)

// -------------------------------------------------------------- arena ----
// Bump allocator over device blocks.  alloc() hands out 256-byte aligned
// slices and adds a block when the current one is full (only while a
// workspace warms up); reset() releases every slice and, if the previous
// pass needed several blocks, replaces them by one block of the total size,
// so steady-state passes make no device allocation at all.  reset() must
// only be called when no kernel of the owner is in flight.
struct Arena {
    struct Block {
        char *base;
        size_t cap;
    };
    std::vector<Block> blocks;
    size_t cur = 0;   // index of the block being filled
    size_t top = 0;   // bytes used in blocks[cur]
    size_t used = 0;  // bytes handed out in this pass
    size_t peak = 0;  // largest pass so far

    int reset();
    bool reset_frees() const { return blocks.size() > 1; }  // reset() would free blocks
    void *alloc_bytes(size_t bytes);  // nullptr + error on failure
    // allocation state: a launch sequence recorded from state m0 to m1 can
    // be replayed (same slices) when the arena is at m0 again, by jump(m1)
    struct Mark {
        size_t nblocks = 0, cur = 0, top = 0, used = 0;
        const char *base0 = nullptr;
    };
    Mark mark() const;
    bool at(const Mark &m) const;
    void jump(const Mark &from, const Mark &to);
    template <typename T>
    T *alloc(size_t count) {
        return reinterpret_cast<T *>(alloc_bytes(count * sizeof(T)));
    }
    void release();
    size_t capacity() const;
};

#define PHF_ALLOC(ptr, arena, T, count)                                      \
    do {                                                                     \
        (ptr) = (arena).template alloc<T>(count);                            \
        if (!(ptr)) return PHF_ERR_NOMEM;                                    \
    } while (0)

// Pinned host staging buffer (grown on demand).  Copies that read or write
// it are asynchronous, so the host neither writes into it nor frees it
// while one may still be pending: mark() records an event after the copies
// enqueued on a stream, and ensure() (before every host use) and release()
// wait for that event first.
struct Pinned {
    char *ptr = nullptr;
    size_t cap = 0;
    hipEvent_t last = nullptr;
    bool pending = false;
    int ensure(size_t bytes);
    int mark(hipStream_t s);
    void release();
};

// ---------------------------------------------------------------- lanes --
struct FmmGraphCache;

struct Lane {
    hipStream_t stream = nullptr;
    Arena arena;
    Pinned pin;
    hipEvent_t ev[2 * PHF_STAGE_COUNT] = {};
    bool busy = false;
    int index = 0;
    // FMM launch sequences recorded as HIP graphs (calls on resident
    // systems, whose buffers keep their addresses between calls), and the
    // second stream (default priority) the FMM near field runs on while the
    // far field runs on `stream`, with the events that fork and join them
    FmmGraphCache *fmm_graphs = nullptr;
    hipStream_t stream2 = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // start, end of convection, end of a time step (its stage times)
    hipEvent_t step_ev[3] = {};
};

// Waits for all work enqueued on the lane's streams (its stream and, once
// created, its second stream).
int lane_quiesce(Lane *L);
// Arena::reset for a lane: when the reset frees device blocks, the lane's
// streams are waited for first, so no enqueued kernel still uses them.
int reset_arena(Lane *L);

class Context {
  public:
    static Context &get();
    int init();  // idempotent; selects device 0 of this process
    int device() const { return dev_; }
    const hipDeviceProp_t &props() const { return props_; }
    int acquire(Lane **out);
    void release(Lane *l);

  private:
    std::mutex mu_;
    std::condition_variable cv_;
    bool ready_ = false;
    int dev_ = -1;
    hipDeviceProp_t props_{};
    std::vector<Lane *> lanes_;
};

// RAII lane borrow.
struct LaneGuard {
    Lane *lane = nullptr;
    int rc = PHF_OK;
    LaneGuard() { rc = Context::get().acquire(&lane); }
    ~LaneGuard() {
        if (lane) Context::get().release(lane);
    }
};

// Per-stage timing with HIP events on the lane's stream; mask selects the
// stages that are timed (bit s for stage s).
struct StageTimer {
    Lane *lane;
    bool on;
    uint32_t mask = ~0u;
    StageTimer(Lane *l, bool enabled) : lane(l), on(enabled) {}
    bool timed(int stage) const { return on && ((mask >> stage) & 1u); }
    void begin(int stage) {
        if (timed(stage)) (void)hipEventRecord(lane->ev[2 * stage], lane->stream);
    }
    void end(int stage) {
        if (timed(stage)) (void)hipEventRecord(lane->ev[2 * stage + 1], lane->stream);
    }
    // the same on another stream of the lane
    void begin_on(int stage, hipStream_t s) {
        if (timed(stage)) (void)hipEventRecord(lane->ev[2 * stage], s);
    }
    void end_on(int stage, hipStream_t s) {
        if (timed(stage)) (void)hipEventRecord(lane->ev[2 * stage + 1], s);
    }
};

// Thread-local record of the stage timings of the last call made by the
// calling thread (filled after the call's stream has been synchronised).
void record_stage_times(Lane *lane, const bool *used);
void get_stage_times(double *ms_out);
void reset_stage_times();
bool timing_enabled();
void set_timing(bool on);

// Looks up (once per device, with that device current) the device address
// of the file-scope scratch array used by phf_vv2par_gauss_noalloc.
int direct_scratch_setup(int dev);

// ---------------------------------------------------------- validation ---
int check_host_array(const double *p, int64_t n, const char *name,
                     bool finite = true);

// ------------------------------------------------------ launch helpers ---
inline unsigned div_up(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace phf
