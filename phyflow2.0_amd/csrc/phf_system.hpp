// phf_system.hpp -- device-resident blob systems (shared by the system,
// step, batch and serve translation units).
#pragma once

#include <mutex>

#include "phf_fmm.hpp"

namespace phf {

struct System {
    std::mutex mu;
    Lane lane;  // private stream, arena, pinned staging and events
    int64_t n = 0, cap = 0;
    double *x = nullptr, *y = nullptr, *w = nullptr;  // resident blobs
    double *vx = nullptr, *vy = nullptr;              // last velocity
    bool has_velocity = false;
};

int sys_reserve(System *s, int64_t n);
int check_vel_params(const phf_vel_params *p);
int velocity_device(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                    const double *xt, const double *yt, int64_t nt, bool same,
                    const phf_vel_params &p, double *vx, double *vy, StageTimer &tm, bool *used,
                    FmmInfo *fi);
int system_velocity_device(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                           const phf_vel_params &p, double *vx, double *vy, StageTimer &tm,
                           bool *used, FmmInfo *fi);

}  // namespace phf
