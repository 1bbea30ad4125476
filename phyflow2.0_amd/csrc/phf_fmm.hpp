// phf_fmm.hpp -- FMM driver interface (device pointers, one stream).
#pragma once

#include "phf_tree.hpp"

#include <functional>

namespace phf {

struct FmmArgs {
    const double *xb, *yb, *wb;  // sources (device, input order)
    int64_t nb;
    const double *xt, *yt;       // targets (device, input order)
    int64_t nt;
    bool same;                   // targets are the sources
    double xopt, cutoff;         // Oseen smoother (fmm2d_py.cpp:61)
    int p, ws, prec, Lhint;      // order, well-separation, precision, leaf level
    double leaf_side = 0.0;      // domain choice (choose_domain_side)
    double *qr, *qi;             // device outputs in target input order
    bool velocity_out = false;   // qr, qi receive (vx, vy) = (-qi, -qr)/(2 pi)
};

struct FmmInfo {
    int L = 0, p = 0, nnear = 0;
    double D = 0.0;
    bool want_counts = false;                // fill d_counts = {P2P pairs, M2L links}
    bool graph = false;                      // record / replay the launches as a HIP graph
    bool no_speculation = false;             // wait for the bounding box before replaying
    // replayed graph: after the stream is synchronised, false when the
    // points needed another leaf level or domain side than the recording:
    // the results are then invalid and the call must be made again with
    // no_speculation
    std::function<bool()> domain_ok;
    unsigned long long *d_counts = nullptr;  // device, arena-owned
    // an evaluation run eagerly whose targets were its sources: the points
    // in the tree's Morton order (arena-owned, valid until the lane's arena
    // is reset): sorted float32 records whose last component holds the
    // input index bits (float32 path), or sorted position -> input index
    // (float64 path); sorted_n points.  Null after a graph replay.
    const float4 *sorted_pay = nullptr;
    const int32_t *sorted_perm = nullptr;
    int64_t sorted_n = 0;
};

int fmm_order_from_tol(double tol);
int fmm_level(int64_t nb, int leaf, int ndirect);
// With info->graph set (and no work counts) the launch sequence (bounding
// box through L2L) is recorded as HIP graphs in the lane's cache and
// replayed on later calls with the same arguments (the stage then timed
// is PHF_STAGE_TREE_FAR); see FmmInfo::domain_ok.  The near field runs on
// the lane's stream while the far field runs on the lane's second stream;
// the lane's stream waits for the far field before the call's last kernel.
int fmm_device(Lane *L, const FmmArgs &a, StageTimer &tm, bool *used, FmmInfo *info);
// the lane's recorded graphs, second stream and its events
void fmm_graphs_release(Lane *L);
// float32 source ranges per leaf padded to multiples of 4: sp[b] (nleaf + 1
// entries) from the leaf offsets; cnt, sums: workspaces (capturable).
int pad_source_ranges(Lane *L, const int32_t *start, int64_t nleaf, uint32_t *cnt, uint32_t *sp,
                      uint32_t *sums);
// split[0..8]: leaf ranges [split[k], split[k+1]) of the 8 XCDs holding
// equal numbers of points of the leaf offsets start[0..nleaf] (capturable).
int xcd_split(Lane *L, const int32_t *start, int64_t nleaf, int32_t *split);
// zero-weight entries in the pads of a padded float32 source array
int zero_source_pads(Lane *L, const int32_t *start, const int32_t *sp, int64_t nleaf, float4 *src);
int q_to_velocity_device(Lane *L, const double *qr, const double *qi, int64_t n, double *vx,
                         double *vy);

}  // namespace phf
