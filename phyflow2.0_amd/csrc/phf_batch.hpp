// phf_batch.hpp -- segment-table batched FMM over many blob systems.
#pragma once

#include <vector>

#include "phf_fmm.hpp"

namespace phf {

// per-system description (host-built once per batch, read by the kernels)
struct BatchSys {
    int64_t off;       // first point in the packed arrays
    int32_t n, L;      // points, leaf level
    int64_t leafbase;  // first global leaf
    double cutoff, c2, inv_ks, scale;    // Oseen smoother (float64)
    float c2_f, nlog2e_ks_f, scale_f;    // float32 forms
    int32_t nocut;                       // float32 cutoff branch is a no-op
};

// per-system square domain, written on the device every evaluation
struct BatchDom {
    double x0, y0, D, s, inv_s;  // origin, side, leaf side, 2^L / D
};

// a run of at most kBatchChunk consecutive points of one system: per-system
// reductions (bounding boxes) go over chunks in parallel, then per system
struct BatchChunk {
    int64_t beg, end;
    int32_t sys, pad;
};
constexpr int64_t kBatchChunk = 4096;

struct BatchPlan {
    int S = 0, p = 0, ws = 1, prec = 0, Lmax = 0;
    int64_t N = 0, nleaf = 0, nbox = 0;
    bool nocut_all = true;
    std::vector<BatchSys> sys;
    std::vector<int64_t> poff, loff, lvl;  // lvl: (Lmax+1) x (S+1), level-major box bases
    std::vector<BatchChunk> chunks;        // in system order
    std::vector<int32_t> cfirst;           // first chunk of each system (S + 1)
};

struct BatchBuffers {
    BatchSys *sys = nullptr;
    BatchDom *dom = nullptr;
    int32_t *split = nullptr;  // leaf ranges of the 8 XCDs (xcd_split)
    int64_t *poff = nullptr, *loff = nullptr, *lvl = nullptr;
    uint32_t *k0 = nullptr, *k1 = nullptr;
    int32_t *v0 = nullptr, *v1 = nullptr;
    double4 *pk = nullptr;  // (x, y, w, 0) per point in input order
    int32_t *start = nullptr;
    char *src = nullptr, *M = nullptr, *B = nullptr;
    char *m2m = nullptr, *l2l = nullptr, *bin = nullptr, *U = nullptr;
    uint32_t *hist = nullptr, *offs = nullptr;
    uint32_t *pcnt = nullptr, *psp = nullptr, *psums = nullptr;  // padded float32 ranges
    int2 *win = nullptr;  // near-window offsets sorted by box gap
    int32_t *flag = nullptr;
    BatchChunk *chunks = nullptr;
    int32_t *cfirst = nullptr;
    char *cpart = nullptr;  // per-chunk partial reductions (32 bytes each)
};

// ---- batched convection + M4' remesh (phf_batch_remesh.hip) ----
struct BatchRm {
    double h, x0, y0, x0h, y0h;  // lattice spacing, origin, origin / h
    int64_t cbase, ccap;         // cell-key range [cbase, cbase + ccap)
};

struct BatchRmDyn {  // written on the device every step
    long long imin, jmin, bi, bj;  // base-cell range
    int32_t ok, pad;               // 0: touched range exceeds ccap
};

struct BatchRemeshPlan {
    phf_batch_remesh_params p{};
    int64_t C = 0;  // total cell capacity
    int key_bits = 1;
    std::vector<BatchRm> rm;
    std::vector<int64_t> cb;  // cbase per system
    std::vector<int32_t> bsys;  // system of each workgroup of cells
};

struct BatchRemeshBuffers {
    Arena arena;
    BatchRm *rm = nullptr;
    BatchRmDyn *dyn = nullptr;
    int64_t *cb = nullptr, *mo = nullptr;  // mo: (offset, count) per system
    int32_t *bsys = nullptr;               // system of each workgroup of cells
    double *xn = nullptr, *yn = nullptr;
    double3 *sb = nullptr;
    // blob range [x, y) of cell k in the sorted order (both 0 for an empty
    // cell)
    int2 *crange = nullptr;
    uint32_t *flag = nullptr, *slot = nullptr, *sums = nullptr;
    double *xo = nullptr, *yo = nullptr, *wo = nullptr;
};

int64_t batch_cell_capacity(int64_t n, double cell_factor);
int batch_remesh_plan(const BatchPlan &P, const double *sigma, const phf_batch_remesh_params &rp,
                      BatchRemeshPlan *R);
int batch_remesh_alloc(const BatchPlan &P, const BatchRemeshPlan &R, BatchRemeshBuffers *Bf);
// Enqueue convection + remesh after batch_enqueue (reuses its sort buffers;
// capturable).
int batch_remesh_enqueue(Lane *L, const BatchPlan &P, BatchBuffers &Bv, const BatchRemeshPlan &R,
                         BatchRemeshBuffers &Bf, const double *x, const double *y,
                         const double *w, const double *vx, const double *vy);

int batch_plan(const int64_t *sizes, const double *sigma, int S, int p, int leaf, int ws, int prec,
               double k, BatchPlan *P);
int batch_alloc(Lane *L, const BatchPlan &P, BatchBuffers *Bf);
// Enqueue one evaluation on L->stream (no allocation, no host sync: safe to
// capture).  vx, vy: device, packed input order.
int batch_enqueue(Lane *L, const BatchPlan &P, BatchBuffers &Bf, const double *x, const double *y,
                  const double *w, double *vx, double *vy);

}  // namespace phf
