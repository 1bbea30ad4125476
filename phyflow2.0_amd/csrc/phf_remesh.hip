// phf_remesh.hip -- deterministic M4' remeshing on gfx950.
//
// Replaces Regrid (src/python/blobs/base/regrid/_regridding.py:36-222) and
// m4kernel_1d (src/cpp/blobs/c++/interpolation_kernels_1d/
// _interpolation_kernels_1d.cpp:21-93).
//
// Reference semantics (restated):
//   iL = int64(ceil(((x - x0) - 2h + h/2)/h)) - 1, likewise jL     (:157-158)
//   stencil slot s = 0..15 -> cell (iL + s%4, jL + s/4)              (:170-172)
//   value = (M4'(ux) * M4'(uy)) * w, u = ((i + 0.5) + x0/h) - x/h    (:183-184)
//   duplicates summed, zero-weight touched cells kept, output in
//   row-major (i, then j) order, x_new = (i h + h/2) + x0            (:200-218)
//
// GPU design (no float atomics, one sort of n keys instead of 16n, no host
// synchronisation before the output count is read):
//   1. each blob's base cell (iL, jL) -> bounds and layout of the touched
//      range, on the device
//   2. stable radix sort of blobs by base-cell key (row-major in range);
//      cell starts from the per-cell populations
//   3. the 16 stencil values of every blob, laid out slot-major
//   4. a cell is touched iff one of its 16 base cells (i-a, j-b),
//      a, b in 0..3, holds a blob; exclusive scan of touched flags gives
//      each output cell its row-major slot
//   5. each touched cell sums its contributions slot-major (s = 0..15),
//      blobs of one base cell in ascending input index: exactly the order
//      of a stable sort of the reference's COO entries (entry e = s*n + b),
//      so the float64 sums are reproducible run to run and match the CPU
//      restatement.
// Kernels over cells run grid-stride over the count the device computed.
// A range too large for the dense layout (far outliers) takes the sparse
// path: occupied base cells and their neighbours as sorted lists (the
// reference's scipy.sparse assembly needs no dense grid either), with the
// same cells, order and sums.
// Index and weight arithmetic runs in float64 with contraction disabled,
// so cell indices and output coordinates are bit-exact to the reference
// formulas; the cube |u|^3 is formed in double-double and rounded once (the
// correctly rounded pow(|u|, 3.0) of the reference).
#include "phf_remesh.hpp"
#include "phf_m4_dev.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

namespace phf {

#pragma clang fp contract(off)

__global__ void m4kernel_1d_kernel(const double *__restrict__ x, int64_t n, double c, double c2,
                                   double *__restrict__ w)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = m4p_dev(x[i], c, c2);
}

// ---------------------------------------------------------- base cells --
constexpr int RM_THREADS = 256;
constexpr int RM_GRID = 2048;  // workgroups of the grid-stride cell kernels

// Geometry of one remesh, computed on the device from the blobs' base cells.
// The dense layout indexes base cells (i - imin) bj + (j - jmin) and touched
// cells (i - imin) nj + (j - jmin), nj = bj + 3; it is used when the touched
// range fits the capacity the buffers were sized for (ok), else the sparse
// path runs from imin .. jmax.
struct RmGeom {
    long long imin, imax, jmin, jmax;  // base-cell bounds
    long long bi, bj, nj;              // dense layout (0 when not ok)
    int64_t nstart;                    // bi bj + 1 cell starts
    int64_t ncell;                     // (bi + 3) nj touched-range cells
    int64_t nscan;                     // ncell + 1
    int64_t m;                         // output cells (set by the gather)
    int32_t ok;
    int32_t bits;                      // key bits of the base-cell index
    int32_t crowded;                   // a base cell holds more than kCellSortMax blobs
    int32_t pad;
};

// Base cells with at most this many blobs have their blobs put in input
// order by one thread (insertion sort); a more crowded cell sends the call
// to the radix sort instead.
constexpr int kCellSortMax = 32;

// Touched-range capacity of the dense layout for n blobs.
static int64_t dense_cells_cap(int64_t n)
{
    return std::min<int64_t>(64 * n + (int64_t(1) << 20), (int64_t(1) << 31) - 2);
}

static int bits_for(uint64_t count)
{
    int b = 1;
    while (b < 32 && (uint64_t(1) << b) < count) ++b;
    return b;
}

// per-block min/max of (iL, jL): longlong4 {imin, imax, jmin, jmax}
__global__ __launch_bounds__(RM_THREADS) void base_bounds_partial(
    const double *__restrict__ x, const double *__restrict__ y, int64_t n, double x0, double y0,
    double h, longlong4 *__restrict__ part)
{
    __shared__ longlong4 sh[RM_THREADS];
    long long imin = LLONG_MAX, imax = LLONG_MIN, jmin = LLONG_MAX, jmax = LLONG_MIN;
    for (int64_t b = (int64_t)blockIdx.x * RM_THREADS + threadIdx.x; b < n;
         b += (int64_t)gridDim.x * RM_THREADS) {
        const long long i = base_index(x[b], x0, h), j = base_index(y[b], y0, h);
        imin = i < imin ? i : imin;
        imax = i > imax ? i : imax;
        jmin = j < jmin ? j : jmin;
        jmax = j > jmax ? j : jmax;
    }
    sh[threadIdx.x] = make_longlong4(imin, imax, jmin, jmax);
    __syncthreads();
    for (int s = RM_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            longlong4 a = sh[threadIdx.x], c = sh[threadIdx.x + s];
            sh[threadIdx.x] = make_longlong4(a.x < c.x ? a.x : c.x, a.y > c.y ? a.y : c.y,
                                             a.z < c.z ? a.z : c.z, a.w > c.w ? a.w : c.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Reduces the partial bounds and lays out the dense range (capacity ccap
// touched cells).
__global__ __launch_bounds__(RM_THREADS) void base_bounds_final(const longlong4 *__restrict__ part,
                                                                int np, int64_t ccap,
                                                                RmGeom *__restrict__ g)
{
    __shared__ longlong4 sh[RM_THREADS];
    longlong4 acc = make_longlong4(LLONG_MAX, LLONG_MIN, LLONG_MAX, LLONG_MIN);
    for (int i = threadIdx.x; i < np; i += RM_THREADS) {
        longlong4 c = part[i];
        acc = make_longlong4(acc.x < c.x ? acc.x : c.x, acc.y > c.y ? acc.y : c.y,
                             acc.z < c.z ? acc.z : c.z, acc.w > c.w ? acc.w : c.w);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = RM_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            longlong4 a = sh[threadIdx.x], c = sh[threadIdx.x + s];
            sh[threadIdx.x] = make_longlong4(a.x < c.x ? a.x : c.x, a.y > c.y ? a.y : c.y,
                                             a.z < c.z ? a.z : c.z, a.w > c.w ? a.w : c.w);
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const longlong4 b = sh[0];
    RmGeom r;
    r.imin = b.x;
    r.imax = b.y;
    r.jmin = b.z;
    r.jmax = b.w;
    // spans in float64 first: far outliers can take them past 64 bits
    const double si = (double)b.y - (double)b.x + 1.0, sj = (double)b.w - (double)b.z + 1.0;
    r.ok = si >= 1.0 && sj >= 1.0 && (si + 3.0) * (sj + 3.0) <= (double)ccap;
    r.bi = r.bj = r.nj = 0;
    r.nstart = r.ncell = r.nscan = 0;
    r.bits = 0;
    r.crowded = 0;
    r.pad = 0;
    r.m = -1;
    if (r.ok) {
        r.bi = b.y - b.x + 1;
        r.bj = b.w - b.z + 1;
        r.nj = r.bj + 3;
        r.nstart = r.bi * r.bj + 1;
        r.ncell = (r.bi + 3) * r.nj;
        r.nscan = r.ncell + 1;
        int k = 1;
        while (k < 31 && (1ll << k) < r.bi * r.bj) ++k;
        r.bits = k;
    }
    *g = r;
}

// v[0 .. *count) = 0 (grid-stride)
__global__ __launch_bounds__(RM_THREADS) void zero_u32_dn(uint32_t *__restrict__ v,
                                                          const int64_t *__restrict__ count)
{
    const int64_t c = *count;
    for (int64_t k = (int64_t)blockIdx.x * RM_THREADS + threadIdx.x; k < c;
         k += (int64_t)gridDim.x * RM_THREADS)
        v[k] = 0u;
}

// key = (iL - imin) bj + (jL - jmin), the blob's base cell in the dense
// range, and the population of each base cell
__global__ void base_keys(const double *__restrict__ x, const double *__restrict__ y, int64_t n,
                          double x0, double y0, double h, const RmGeom *__restrict__ g,
                          uint32_t *__restrict__ keys, int32_t *__restrict__ vals,
                          uint32_t *__restrict__ count)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    uint32_t key = 0u;
    if (g->ok) {
        const long long i = base_index(x[b], x0, h) - g->imin, j = base_index(y[b], y0, h) - g->jmin;
        key = (uint32_t)(i * g->bj + j);
        atomicAdd(count + key, 1u);
    }
    keys[b] = key;
    vals[b] = (int32_t)b;
}

// Counting sort by base cell: blob b takes a place in its cell's range
// [start[key], start[key + 1]) in arrival order (cursor: zeroed counts) ...
__global__ void cell_scatter(const uint32_t *__restrict__ keys, int64_t n,
                             const RmGeom *__restrict__ g, const uint32_t *__restrict__ start,
                             uint32_t *__restrict__ cursor, int32_t *__restrict__ perm)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n || !g->ok) return;
    const uint32_t key = keys[b];
    perm[start[key] + atomicAdd(cursor + key, 1u)] = (int32_t)b;
}

// ... then each cell's blobs in ascending input index: the order of a
// stable sort by base cell
__global__ __launch_bounds__(RM_THREADS) void cell_order(const uint32_t *__restrict__ start,
                                                         RmGeom *__restrict__ g,
                                                         int32_t *__restrict__ perm)
{
    if (!g->ok) return;
    const int64_t nb = g->nstart - 1;
    for (int64_t k = (int64_t)blockIdx.x * RM_THREADS + threadIdx.x; k < nb;
         k += (int64_t)gridDim.x * RM_THREADS) {
        const uint32_t lo = start[k], hi = start[k + 1];
        if (hi - lo < 2u) continue;
        if (hi - lo > (uint32_t)kCellSortMax) {
            g->crowded = 1;
            continue;
        }
        for (uint32_t i = lo + 1; i < hi; ++i) {
            const int32_t v = perm[i];
            uint32_t j = i;
            while (j > lo && perm[j - 1] > v) {
                perm[j] = perm[j - 1];
                --j;
            }
            perm[j] = v;
        }
    }
}

// The 16 stencil values of each blob in base-cell order: V[s n + q] for
// sorted blob q = perm^-1(b) and slot s = a + 4 b'.  M4' (KIND 0):
// (M4'(ux) M4'(uy)) w with u = ((i + 0.5) + x0/h) - x/h, i = iL + a
// (_regridding.py:170-184); Tutty (KIND 1): (F[a] G[b']) w with F, G the
// weights about the node right of the base cell (_evolution.py:385-405).
// Each value feeds exactly one output cell.
template <int KIND>
__global__ __launch_bounds__(RM_THREADS) void stencil_values(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ w,
    const int32_t *__restrict__ perm, int64_t n, double x0, double y0, double h, double x0h,
    double y0h, double c, double c2, double *__restrict__ V)
{
    const int64_t q = (int64_t)blockIdx.x * RM_THREADS + threadIdx.x;
    if (q >= n) return;
    const int32_t b = perm[q];
    const double xb = x[b], yb = y[b], wb = w[b];
    const long long iL = base_index(xb, x0, h), jL = base_index(yb, y0, h);
    double fx[4], fy[4];
    if (KIND == 1) {
        const double hh = 0.5 * h;
        const double xc = ((double)(iL + 1) * h + hh) + x0;
        const double yc = ((double)(jL + 1) * h + hh) + y0;
        const double dx = (xb - xc) / h, dy = (yb - yc) / h;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            fx[a] = tutty_weight(a, dx, c2);
            fy[a] = tutty_weight(a, dy, c2);
        }
    } else {
        const double px = xb / h, py = yb / h;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            fx[a] = m4p_dev((((double)(iL + a) + 0.5) + x0h) - px, c, c2);
            fy[a] = m4p_dev((((double)(jL + a) + 0.5) + y0h) - py, c, c2);
        }
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) V[(int64_t)s * n + q] = (fx[s & 3] * fy[s >> 2]) * wb;
}

// Starts of the 16 base cells (i - a, j - b) of touched cell (i, j): st[a][e]
// is the start of base cell (i - a, j - 3 + e), e = 0..4 (e = 4 closes
// column j), 0 outside the range.
__device__ __forceinline__ void base_starts(const int32_t *__restrict__ start, long long i,
                                            long long j, long long bi, long long bj,
                                            int32_t st[4][5], bool rowok[4])
{
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const long long ii = i - a;
        rowok[a] = ii >= 0 && ii < bi;
#pragma unroll
        for (int e = 0; e < 5; ++e) {
            const long long jj = j - 3 + e;
            st[a][e] = (rowok[a] && jj >= 0 && jj <= bj) ? start[ii * bj + jj] : 0;
        }
    }
}

// touched flag of every cell of the dense range (grid-stride), and a
// closing 0 so that the exclusive scan ends in the output count
__global__ __launch_bounds__(RM_THREADS) void touched_cells(const int32_t *__restrict__ start,
                                                            const RmGeom *__restrict__ g,
                                                            uint32_t *__restrict__ flag)
{
    if (!g->ok) return;
    const int64_t ncell = g->ncell;
    const long long bi = g->bi, bj = g->bj, nj = g->nj;
    for (int64_t cidx = (int64_t)blockIdx.x * RM_THREADS + threadIdx.x; cidx <= ncell;
         cidx += (int64_t)gridDim.x * RM_THREADS) {
        if (cidx == ncell) {
            flag[cidx] = 0u;
            continue;
        }
        long long i, j;
        divmod_cells(cidx, nj, i, j);
        int32_t st[4][5];
        bool rowok[4];
        base_starts(start, i, j, bi, bj, st, rowok);
        uint32_t t = 0u;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
            const int a = s & 3, b = s >> 2;
            const long long jj = j - b;
            t |= (rowok[a] && jj >= 0 && jj < bj && st[a][4 - b] > st[a][3 - b]) ? 1u : 0u;
        }
        flag[cidx] = t;
    }
}

// Sum of touched cell (i, j) slot-major (s = 0..15), blobs of one base cell
// in ascending input index: the order of a stable sort of the reference's
// COO entries (entry e = s n + b).
__device__ __forceinline__ double cell_sum_dense(const int32_t st[4][5], const bool rowok[4],
                                                 long long j, long long bj,
                                                 const double *__restrict__ V, int64_t n)
{
    double acc = 0.0;
    bool first = true;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int a = s & 3, b = s >> 2;
        const long long jj = j - b;
        if (!(rowok[a] && jj >= 0 && jj < bj)) continue;
        const double *Vs = V + (int64_t)s * n;
        for (int32_t q = st[a][3 - b]; q < st[a][4 - b]; ++q) {
            const double v = Vs[q];
            acc = first ? v : acc + v;
            first = false;
        }
    }
    return acc;
}

__global__ __launch_bounds__(RM_THREADS) void gather_cells(
    const int32_t *__restrict__ start, const double *__restrict__ V, int64_t n,
    const uint32_t *__restrict__ slot, RmGeom *__restrict__ g, double h, double x0, double y0,
    double *__restrict__ xo, double *__restrict__ yo, double *__restrict__ wo,
    int64_t *__restrict__ io, int64_t *__restrict__ jo)
{
    if (!g->ok) return;
    const int64_t ncell = g->ncell;
    const long long bi = g->bi, bj = g->bj, nj = g->nj, imin = g->imin, jmin = g->jmin;
    const int64_t t0 = (int64_t)blockIdx.x * RM_THREADS + threadIdx.x;
    if (t0 == 0) g->m = slot[ncell];
    const double hh = 0.5 * h;
    for (int64_t cidx = t0; cidx < ncell; cidx += (int64_t)gridDim.x * RM_THREADS) {
        const uint32_t o = slot[cidx];
        if (slot[cidx + 1] == o) continue;  // untouched
        long long i, j;
        divmod_cells(cidx, nj, i, j);
        int32_t st[4][5];
        bool rowok[4];
        base_starts(start, i, j, bi, bj, st, rowok);
        const long long I = i + imin, J = j + jmin;  // absolute cell indices
        wo[o] = cell_sum_dense(st, rowok, j, bj, V, n);
        xo[o] = ((double)I * h + hh) + x0;
        yo[o] = ((double)J * h + hh) + y0;
        if (io) io[o] = I;
        if (jo) jo[o] = J;
    }
}

// ------------------------------------------------------ sparse fallback --
// Used when the blobs' bounding range is too large for the dense layout
// (distant outliers, long thin wakes): only occupied base cells and their
// 16 neighbours are kept, as sorted lists, which gives the same cells in
// the same order and the same sums.

__global__ void sparse_base_ij(const double *__restrict__ x, const double *__restrict__ y,
                               int64_t n, double x0, double y0, double h, long long imin,
                               long long jmin, uint32_t *__restrict__ bi_, uint32_t *__restrict__ bj_,
                               int32_t *__restrict__ vals)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    bi_[b] = (uint32_t)(base_index(x[b], x0, h) - imin);
    bj_[b] = (uint32_t)(base_index(y[b], y0, h) - jmin);
    vals[b] = (int32_t)b;
}

// out[q] = in[idx[q]]
__global__ void gather_u32(const uint32_t *__restrict__ in, const int32_t *__restrict__ idx,
                           int64_t n, uint32_t *__restrict__ out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) out[q] = in[idx[q]];
}

// flag[q] = 1 where (a, b)[q] starts a run of equal pairs; flag[n] = 0
__global__ void run_heads(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b, int64_t n,
                          uint32_t *__restrict__ flag)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q > n) return;
    flag[q] = q == n ? 0u : (q == 0 || a[q] != a[q - 1] || b[q] != b[q - 1]) ? 1u : 0u;
}

// unique pairs as keys (a << 32 | b) with their first position
__global__ void run_compact(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                            int64_t n, const uint32_t *__restrict__ pos,
                            uint64_t *__restrict__ ukey, int32_t *__restrict__ ustart)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q > n) return;
    if (q == n) {
        ustart[pos[n]] = (int32_t)n;
        return;
    }
    if (pos[q + 1] == pos[q]) return;
    ukey[pos[q]] = ((uint64_t)a[q] << 32) | b[q];
    ustart[pos[q]] = (int32_t)q;
}

// the 16 touched cells (i + a, j + b) of each occupied base cell
__global__ void stencil_cells(const uint64_t *__restrict__ ukey, int64_t U,
                              uint32_t *__restrict__ ci, int32_t *__restrict__ cj)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 16 * U) return;
    const int64_t u = e % U;
    const int s = (int)(e / U);
    ci[e] = (uint32_t)(ukey[u] >> 32) + (uint32_t)(s & 3);
    cj[e] = (int32_t)((uint32_t)ukey[u] + (uint32_t)(s >> 2));
}

__device__ __forceinline__ int64_t find_key(const uint64_t *__restrict__ k, int64_t U, uint64_t v)
{
    int64_t lo = 0, hi = U;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (k[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return (lo < U && k[lo] == v) ? lo : -1;
}

__global__ __launch_bounds__(RM_THREADS) void gather_sparse(
    const uint64_t *__restrict__ tkey, int64_t M, const uint64_t *__restrict__ ukey,
    const int32_t *__restrict__ ustart, int64_t U, const double *__restrict__ V, int64_t n,
    long long imin, long long jmin, double h, double x0, double y0, double *__restrict__ xo,
    double *__restrict__ yo, double *__restrict__ wo, int64_t *__restrict__ io,
    int64_t *__restrict__ jo)
{
    const int64_t t = (int64_t)blockIdx.x * RM_THREADS + threadIdx.x;
    if (t >= M) return;
    const long long i = (long long)(tkey[t] >> 32), j = (long long)(uint32_t)tkey[t];
    double acc = 0.0;
    bool first = true;
    for (int s = 0; s < 16; ++s) {
        const long long ii = i - (s & 3), jj = j - (s >> 2);
        if (ii < 0 || jj < 0) continue;
        const int64_t u = find_key(ukey, U, ((uint64_t)ii << 32) | (uint64_t)jj);
        if (u < 0) continue;
        const double *Vs = V + (int64_t)s * n;
        for (int32_t q = ustart[u]; q < ustart[u + 1]; ++q) {
            const double v = Vs[q];
            acc = first ? v : acc + v;
            first = false;
        }
    }
    const long long I = i + imin, J = j + jmin;
    const double hh = 0.5 * h;
    wo[t] = acc;
    xo[t] = ((double)I * h + hh) + x0;
    yo[t] = ((double)J * h + hh) + y0;
    if (io) io[t] = I;
    if (jo) jo[t] = J;
}

static int launch_stencil_values(Lane *L, int kind, const double *x, const double *y,
                                 const double *w, const int32_t *perm, int64_t n, double x0,
                                 double y0, double h, double c, double *V)
{
    if (kind == 1) {
        const double r = c / h;  // c carries hNu = sqrt(nu dt)
        stencil_values<1><<<div_up(n, RM_THREADS), RM_THREADS, 0, L->stream>>>(
            x, y, w, perm, n, x0, y0, h, 0.0, 0.0, 0.0, r * r, V);
    } else {
        stencil_values<0><<<div_up(n, RM_THREADS), RM_THREADS, 0, L->stream>>>(
            x, y, w, perm, n, x0, y0, h, x0 / h, y0 / h, c, c * c, V);
    }
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

static int read_count(Lane *L, const uint32_t *d, int64_t *out)
{
    uint32_t v = 0;
    PHF_HIP(hipMemcpyAsync(&v, d, sizeof(v), hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    *out = v;
    return PHF_OK;
}

static int remesh_sparse(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                         double h, double x0, double y0, double c, int kind, const RmGeom &hg,
                         RemeshOut *out)
{
    const double si = (double)hg.imax - (double)hg.imin + 1.0;
    const double sj = (double)hg.jmax - (double)hg.jmin + 1.0;
    if (!(si >= 1.0 && sj >= 1.0 && si + 3.0 < 2147483647.0 && sj + 3.0 < 2147483647.0)) {
        set_error("remesh: the blobs span %.3g x %.3g cells (at most 2^31 - 4 per axis)", si, sj);
        return PHF_ERR_UNSUPPORTED;
    }
    const uint32_t bi = (uint32_t)si, bj = (uint32_t)sj;
    const size_t nn = (size_t)n;
    // 1. blobs sorted by base cell (i, j): by j, then stably by i
    uint32_t *ki, *kj, *t0, *t1, *t2;
    int32_t *v0, *v1;
    PHF_ALLOC(ki, L->arena, uint32_t, nn);
    PHF_ALLOC(kj, L->arena, uint32_t, nn);
    PHF_ALLOC(t0, L->arena, uint32_t, nn);
    PHF_ALLOC(t1, L->arena, uint32_t, nn);
    PHF_ALLOC(t2, L->arena, uint32_t, nn);
    PHF_ALLOC(v0, L->arena, int32_t, nn);
    PHF_ALLOC(v1, L->arena, int32_t, nn);
    sparse_base_ij<<<div_up(n, 256), 256, 0, L->stream>>>(x, y, n, x0, y0, h, hg.imin, hg.jmin,
                                                          ki, kj, v0);
    PHF_HIP(hipGetLastError());
    PHF_HIP(hipMemcpyAsync(t0, kj, sizeof(uint32_t) * nn, hipMemcpyDeviceToDevice, L->stream));
    uint32_t *ks;
    int32_t *p1, *perm;
    PHF_TRY(device_radix_sort(L, t0, v0, t1, v1, n, bits_for(bj), &ks, &p1));
    uint32_t *ka = ks == t0 ? t1 : t0;  // the free key buffer
    gather_u32<<<div_up(n, 256), 256, 0, L->stream>>>(ki, p1, n, ka);
    PHF_TRY(device_radix_sort(L, ka, p1, ks, p1 == v0 ? v1 : v0, n, bits_for(bi), &ks, &perm));
    uint32_t *is = ks, *js = t2;
    gather_u32<<<div_up(n, 256), 256, 0, L->stream>>>(kj, perm, n, js);
    PHF_HIP(hipGetLastError());
    // 2. occupied base cells: keys (i << 32 | j) and blob ranges
    uint32_t *flag, *pos, *sums;
    PHF_ALLOC(flag, L->arena, uint32_t, nn + 1);
    PHF_ALLOC(pos, L->arena, uint32_t, nn + 1);
    PHF_ALLOC(sums, L->arena, uint32_t, (size_t)scan_sums_size(n + 1));
    run_heads<<<div_up(n + 1, 256), 256, 0, L->stream>>>(is, js, n, flag);
    PHF_TRY(device_exclusive_scan_ws(L, flag, n + 1, pos, sums));
    int64_t U = 0;
    PHF_TRY(read_count(L, pos + n, &U));
    uint64_t *ukey;
    int32_t *ustart;
    PHF_ALLOC(ukey, L->arena, uint64_t, (size_t)U);
    PHF_ALLOC(ustart, L->arena, int32_t, (size_t)U + 1);
    run_compact<<<div_up(n + 1, 256), 256, 0, L->stream>>>(is, js, n, pos, ukey, ustart);
    PHF_HIP(hipGetLastError());
    // 3. touched cells: the 16 neighbours of each occupied base cell, sorted
    //    row-major (by j, then stably by i) and made unique
    const int64_t E = 16 * U;
    const size_t ne = (size_t)E;
    uint32_t *ci, *ck0, *ck1;
    int32_t *cj, *cv0, *cv1;
    PHF_ALLOC(ci, L->arena, uint32_t, ne);
    PHF_ALLOC(cj, L->arena, int32_t, ne);
    PHF_ALLOC(ck0, L->arena, uint32_t, ne);
    PHF_ALLOC(ck1, L->arena, uint32_t, ne);
    PHF_ALLOC(cv0, L->arena, int32_t, ne);
    PHF_ALLOC(cv1, L->arena, int32_t, ne);
    stencil_cells<<<div_up(E, 256), 256, 0, L->stream>>>(ukey, U, ci, cj);
    PHF_HIP(hipGetLastError());
    // sort 1: key j, value i
    PHF_HIP(hipMemcpyAsync(ck0, cj, sizeof(uint32_t) * ne, hipMemcpyDeviceToDevice, L->stream));
    PHF_HIP(hipMemcpyAsync(cv0, ci, sizeof(int32_t) * ne, hipMemcpyDeviceToDevice, L->stream));
    uint32_t *sj_k;
    int32_t *si_v;
    PHF_TRY(device_radix_sort(L, ck0, cv0, ck1, cv1, E, bits_for((uint64_t)bj + 3), &sj_k, &si_v));
    // sort 2: key i, value j
    uint32_t *k2 = sj_k == ck0 ? ck1 : ck0;
    int32_t *v2 = si_v == cv0 ? cv1 : cv0;
    PHF_HIP(hipMemcpyAsync(k2, si_v, sizeof(uint32_t) * ne, hipMemcpyDeviceToDevice, L->stream));
    PHF_HIP(hipMemcpyAsync(v2, sj_k, sizeof(int32_t) * ne, hipMemcpyDeviceToDevice, L->stream));
    uint32_t *ti;
    int32_t *tj;
    PHF_TRY(device_radix_sort(L, k2, v2, sj_k, si_v, E, bits_for((uint64_t)bi + 3), &ti, &tj));
    uint32_t *cflag, *cpos, *csums;
    PHF_ALLOC(cflag, L->arena, uint32_t, ne + 1);
    PHF_ALLOC(cpos, L->arena, uint32_t, ne + 1);
    PHF_ALLOC(csums, L->arena, uint32_t, (size_t)scan_sums_size(E + 1));
    run_heads<<<div_up(E + 1, 256), 256, 0, L->stream>>>(ti, reinterpret_cast<uint32_t *>(tj), E,
                                                         cflag);
    PHF_TRY(device_exclusive_scan_ws(L, cflag, E + 1, cpos, csums));
    int64_t M = 0;
    PHF_TRY(read_count(L, cpos + E, &M));
    uint64_t *tkey;
    int32_t *tstart;
    PHF_ALLOC(tkey, L->arena, uint64_t, (size_t)std::max<int64_t>(M, 1));
    PHF_ALLOC(tstart, L->arena, int32_t, (size_t)M + 1);
    run_compact<<<div_up(E + 1, 256), 256, 0, L->stream>>>(ti, reinterpret_cast<uint32_t *>(tj),
                                                           E, cpos, tkey, tstart);
    PHF_HIP(hipGetLastError());
    // 4. values and sums
    double *V;
    PHF_ALLOC(V, L->arena, double, 16 * nn);
    PHF_TRY(launch_stencil_values(L, kind, x, y, w, perm, n, x0, y0, h, c, V));
    double *xo, *yo, *wo;
    int64_t *io = nullptr, *jo = nullptr;
    const size_t mm = (size_t)std::max<int64_t>(M, 1);
    PHF_ALLOC(xo, L->arena, double, mm);
    PHF_ALLOC(yo, L->arena, double, mm);
    PHF_ALLOC(wo, L->arena, double, mm);
    if (out->want_indices) {
        PHF_ALLOC(io, L->arena, int64_t, mm);
        PHF_ALLOC(jo, L->arena, int64_t, mm);
    }
    gather_sparse<<<div_up(M, RM_THREADS), RM_THREADS, 0, L->stream>>>(
        tkey, M, ukey, ustart, U, V, n, hg.imin, hg.jmin, h, x0, y0, xo, yo, wo, io, jo);
    PHF_HIP(hipGetLastError());
    out->m = M;
    out->x = xo;
    out->y = yo;
    out->w = wo;
    out->i = io;
    out->j = jo;
    out->cells = M;
    out->sparse = true;
    return PHF_OK;
}

// Dense path: no host synchronisation until the output count is read at
// the end (with whatever the caller queued before the call).  The touched
// range is laid out on the device; if it exceeds the capacity the buffers
// were sized for, the sparse path runs instead.
int remesh_device(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                  double sigma, double overlap, double x0, double y0, double c, RemeshOut *out,
                  int kind)
{
    out->m = 0;
    out->sparse = false;
    if (n == 0) return PHF_OK;
    PHF_REQUIRE(n < (int64_t(1) << 31), "remesh of %lld blobs (at most 2^31 - 1)", (long long)n);
    const double h = sigma * overlap;
    const int64_t ccap = dense_cells_cap(n);
    const size_t nn = (size_t)n;
    // 1. bounds and layout of the base cells
    const int nblk = (int)std::min<int64_t>(1024, (n + RM_THREADS * 8 - 1) / (RM_THREADS * 8));
    longlong4 *part;
    RmGeom *g;
    PHF_ALLOC(part, L->arena, longlong4, (size_t)nblk);
    PHF_ALLOC(g, L->arena, RmGeom, 1);
    base_bounds_partial<<<nblk, RM_THREADS, 0, L->stream>>>(x, y, n, x0, y0, h, part);
    base_bounds_final<<<1, RM_THREADS, 0, L->stream>>>(part, nblk, ccap, g);
    PHF_HIP(hipGetLastError());
    // 2. blobs ordered by base cell (counting sort; the radix sort when a
    //    cell is crowded); cell starts from the cell populations
    uint32_t *k0, *cnt, *sums, *ustart;
    int32_t *v0;
    PHF_ALLOC(k0, L->arena, uint32_t, nn);
    PHF_ALLOC(v0, L->arena, int32_t, nn);
    PHF_ALLOC(cnt, L->arena, uint32_t, (size_t)ccap + 1);
    PHF_ALLOC(ustart, L->arena, uint32_t, (size_t)ccap + 1);
    PHF_ALLOC(sums, L->arena, uint32_t, (size_t)scan_sums_size(ccap + 1));
    zero_u32_dn<<<RM_GRID, RM_THREADS, 0, L->stream>>>(cnt, &g->nstart);
    base_keys<<<div_up(n, 256), 256, 0, L->stream>>>(x, y, n, x0, y0, h, g, k0, v0, cnt);
    PHF_HIP(hipGetLastError());
    PHF_TRY(device_exclusive_scan_ws(L, cnt, ccap + 1, ustart, sums, &g->nstart));
    const int32_t *start = reinterpret_cast<const int32_t *>(ustart);
    // the touched-cell scan below overwrites slot; until then it is the cursor
    uint32_t *slot;
    PHF_ALLOC(slot, L->arena, uint32_t, (size_t)ccap + 1);
    // outputs sized for the most cells n blobs can touch
    const size_t cap = (size_t)std::min<int64_t>(16 * n, ccap);
    double *V, *xo, *yo, *wo;
    int64_t *io = nullptr, *jo = nullptr;
    PHF_ALLOC(V, L->arena, double, 16 * nn);
    PHF_ALLOC(xo, L->arena, double, cap);
    PHF_ALLOC(yo, L->arena, double, cap);
    PHF_ALLOC(wo, L->arena, double, cap);
    if (out->want_indices) {
        PHF_ALLOC(io, L->arena, int64_t, cap);
        PHF_ALLOC(jo, L->arena, int64_t, cap);
    }
    RmGeom hg;
    for (int radix = 0; radix < 2; ++radix) {
        int32_t *perm = v0;
        if (!radix) {
            int32_t *pc;
            PHF_ALLOC(pc, L->arena, int32_t, nn);
            zero_u32_dn<<<RM_GRID, RM_THREADS, 0, L->stream>>>(slot, &g->nstart);
            cell_scatter<<<div_up(n, 256), 256, 0, L->stream>>>(k0, n, g, ustart, slot, pc);
            cell_order<<<RM_GRID, RM_THREADS, 0, L->stream>>>(ustart, g, pc);
            PHF_HIP(hipGetLastError());
            perm = pc;
        } else {
            uint32_t *k1, *ks;
            int32_t *v1;
            PHF_ALLOC(k1, L->arena, uint32_t, nn);
            PHF_ALLOC(v1, L->arena, int32_t, nn);
            PHF_TRY(device_radix_sort_dbits(L, k0, v0, k1, v1, n, bits_for((uint64_t)ccap),
                                            &g->bits, &ks, &perm));
        }
        // 3. stencil values of every blob
        PHF_TRY(launch_stencil_values(L, kind, x, y, w, perm, n, x0, y0, h, c, V));
        // 4. touched cells and their output slots
        uint32_t *flag = cnt;  // the populations are no longer needed
        touched_cells<<<RM_GRID, RM_THREADS, 0, L->stream>>>(start, g, flag);
        PHF_HIP(hipGetLastError());
        PHF_TRY(device_exclusive_scan_ws(L, flag, ccap + 1, slot, sums, &g->nscan));
        // 5. sums
        gather_cells<<<RM_GRID, RM_THREADS, 0, L->stream>>>(start, V, n, slot, g, h, x0, y0, xo,
                                                            yo, wo, io, jo);
        PHF_HIP(hipGetLastError());
        // the one synchronisation: the layout with the output count, and what
        // the caller asked to read back
        PHF_TRY(L->pin.ensure(sizeof(RmGeom) + out->read_bytes));
        PHF_HIP(hipMemcpyAsync(L->pin.ptr, g, sizeof(RmGeom), hipMemcpyDeviceToHost, L->stream));
        if (out->read_bytes)
            PHF_HIP(hipMemcpyAsync(L->pin.ptr + sizeof(RmGeom), out->read_src, out->read_bytes,
                                   hipMemcpyDeviceToHost, L->stream));
        PHF_TRY(L->pin.mark(L->stream));
        PHF_HIP(hipStreamSynchronize(L->stream));
        std::memcpy(&hg, L->pin.ptr, sizeof(hg));
        if (out->read_bytes)
            std::memcpy(out->read_dst, L->pin.ptr + sizeof(RmGeom), out->read_bytes);
        if (!(hg.ok && hg.crowded)) break;
    }
    if (!hg.ok) return remesh_sparse(L, x, y, w, n, h, x0, y0, c, kind, hg, out);
    PHF_REQUIRE(hg.m >= 0 && hg.m <= (int64_t)cap, "remesh output count %lld out of range",
                (long long)hg.m);
    out->m = hg.m;
    out->x = xo;
    out->y = yo;
    out->w = wo;
    out->i = io;
    out->j = jo;
    out->cells = hg.ncell;
    return PHF_OK;
}

int m4kernel_device(Lane *L, const double *x, int64_t n, double c, double *w)
{
    if (n <= 0) return PHF_OK;
    m4kernel_1d_kernel<<<div_up(n, 256), 256, 0, L->stream>>>(x, n, c, c * c, w);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

// ------------------------------------------------------------ moments --
// Float64 invariants sum w, sum w x, sum w y, sum w (x^2 + y^2): per-block
// tree sums then one block over the partials, fixed order throughout.
__global__ __launch_bounds__(RM_THREADS) void moments_partial(const double *__restrict__ x,
                                                              const double *__restrict__ y,
                                                              const double *__restrict__ w,
                                                              int64_t n, double4 *__restrict__ part)
{
    __shared__ double4 sh[RM_THREADS];
    double a = 0, b = 0, c = 0, d = 0;
    for (int64_t i = (int64_t)blockIdx.x * RM_THREADS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RM_THREADS) {
        const double wi = w[i], xi = x[i], yi = y[i];
        a += wi;
        b += wi * xi;
        c += wi * yi;
        d += wi * (xi * xi + yi * yi);
    }
    sh[threadIdx.x] = make_double4(a, b, c, d);
    __syncthreads();
    for (int s = RM_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            double4 p = sh[threadIdx.x], q = sh[threadIdx.x + s];
            sh[threadIdx.x] = make_double4(p.x + q.x, p.y + q.y, p.z + q.z, p.w + q.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(RM_THREADS) void moments_final(const double4 *__restrict__ part,
                                                            int np, double4 *__restrict__ out)
{
    __shared__ double4 sh[RM_THREADS];
    double4 acc = make_double4(0, 0, 0, 0);
    for (int i = threadIdx.x; i < np; i += RM_THREADS) {
        double4 q = part[i];
        acc = make_double4(acc.x + q.x, acc.y + q.y, acc.z + q.z, acc.w + q.w);
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = RM_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            double4 p = sh[threadIdx.x], q = sh[threadIdx.x + s];
            sh[threadIdx.x] = make_double4(p.x + q.x, p.y + q.y, p.z + q.z, p.w + q.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

int moments_device(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                   double4 *d_out)
{
    const int nblk = 256;
    double4 *part;
    PHF_ALLOC(part, L->arena, double4, nblk);
    moments_partial<<<nblk, RM_THREADS, 0, L->stream>>>(x, y, w, n, part);
    moments_final<<<1, RM_THREADS, 0, L->stream>>>(part, nblk, d_out);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

// ---------------------------------------------------- population control --
// PopulationControl (_regridding.py:227-330): discard |w| < thr, where thr
// starts at wThreshold and is divided by 10 until the discarded total |w|
// falls below wTotalThreshold; kept blobs keep their order.  The losses of
// all candidate thresholds are formed in one pass (float64, fixed order),
// the host picks the threshold, a stable scan compacts.
constexpr int PC_K = 32;

__global__ __launch_bounds__(256) void pc_loss_partial(const double *__restrict__ w, int64_t n,
                                                       const double *__restrict__ thr, int K,
                                                       double *__restrict__ part)
{
    __shared__ double sh[256];
    double acc[PC_K];
#pragma unroll
    for (int k = 0; k < PC_K; ++k) acc[k] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double a = fabs(w[i]);
#pragma unroll
        for (int k = 0; k < PC_K; ++k)
            if (k < K && a < thr[k]) acc[k] += a;
    }
    for (int k = 0; k < K; ++k) {
        sh[threadIdx.x] = acc[k];
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * K + k] = sh[0];
        __syncthreads();
    }
}

__global__ void pc_loss_final(const double *__restrict__ part, int nblk, int K,
                              double *__restrict__ loss)
{
    const int k = threadIdx.x;
    if (k >= K) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * K + k];
    loss[k] = s;
}

__global__ void pc_flags(const double *__restrict__ w, int64_t n, double thr,
                         uint32_t *__restrict__ keep)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keep[i] = (fabs(w[i]) < thr) ? 0u : 1u;
}

__global__ void pc_scatter(const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ w, int64_t n,
                           const uint32_t *__restrict__ keep, const uint32_t *__restrict__ pos,
                           double *__restrict__ xo, double *__restrict__ yo, double *__restrict__ wo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const uint32_t o = pos[i];
    xo[o] = x[i];
    yo[o] = y[i];
    wo[o] = w[i];
}

int population_control_device(Lane *L, const double *x, const double *y, const double *w,
                              int64_t n, double wThreshold, double wTotalThreshold, double *xo,
                              double *yo, double *wo, int64_t *m_out, double *thr_used)
{
    *m_out = 0;
    if (thr_used) *thr_used = wThreshold;
    if (n == 0) return PHF_OK;
    // candidate thresholds by repeated division, as the reference loop does
    std::vector<double> thr(PC_K);
    thr[0] = wThreshold;
    for (int k = 1; k < PC_K; ++k) thr[k] = thr[k - 1] / 10.0;
    const int nblk = (int)std::min<int64_t>(256, (n + 255) / 256);
    double *dthr, *part, *loss;
    PHF_ALLOC(dthr, L->arena, double, PC_K);
    PHF_ALLOC(part, L->arena, double, (size_t)nblk * PC_K);
    PHF_ALLOC(loss, L->arena, double, PC_K);
    PHF_TRY(L->pin.ensure(sizeof(double) * 2 * PC_K));
    double *hp = reinterpret_cast<double *>(L->pin.ptr);
    std::memcpy(hp, thr.data(), sizeof(double) * PC_K);
    PHF_HIP(hipMemcpyAsync(dthr, hp, sizeof(double) * PC_K, hipMemcpyHostToDevice, L->stream));
    pc_loss_partial<<<nblk, 256, 0, L->stream>>>(w, n, dthr, PC_K, part);
    pc_loss_final<<<1, 64, 0, L->stream>>>(part, nblk, PC_K, loss);
    PHF_HIP(hipGetLastError());
    PHF_HIP(hipMemcpyAsync(hp + PC_K, loss, sizeof(double) * PC_K, hipMemcpyDeviceToHost,
                           L->stream));
    PHF_TRY(L->pin.mark(L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    double t = 0.0;  // no acceptable threshold among the candidates: keep all
    for (int k = 0; k < PC_K; ++k)
        if (hp[PC_K + k] < wTotalThreshold) {
            t = thr[k];
            break;
        }
    if (thr_used) *thr_used = t;
    uint32_t *keep, *pos;
    PHF_ALLOC(keep, L->arena, uint32_t, (size_t)n);
    PHF_ALLOC(pos, L->arena, uint32_t, (size_t)n);
    pc_flags<<<div_up(n, 256), 256, 0, L->stream>>>(w, n, t, keep);
    PHF_TRY(device_exclusive_scan(L, keep, n, pos));
    pc_scatter<<<div_up(n, 256), 256, 0, L->stream>>>(x, y, w, n, keep, pos, xo, yo, wo);
    PHF_HIP(hipGetLastError());
    uint32_t tail[2];
    PHF_HIP(hipMemcpyAsync(&tail[0], pos + n - 1, 4, hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipMemcpyAsync(&tail[1], keep + n - 1, 4, hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    *m_out = (int64_t)tail[0] + tail[1];
    return PHF_OK;
}

}  // namespace phf

// ------------------------------------------------------------ C entry ----
using namespace phf;

struct phf_blobset_s {
    int64_t m = 0;
    std::vector<double> x, y, w;
    std::vector<int64_t> i, j;
};

extern "C" {

int phf_m4kernel_1d(const double *x, int64_t n, double c, double *w)
{
    clear_error();
    PHF_REQUIRE(n >= 0, "negative size");
    PHF_REQUIRE(n == 0 || (x && w), "NULL array");
    PHF_REQUIRE(std::isfinite(c), "c must be finite");
    if (n == 0) return PHF_OK;
    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *L = g.lane;
    PHF_TRY(reset_arena(L));
    double *dx, *dw;
    PHF_TRY(upload(L, x, n, &dx));
    PHF_ALLOC(dw, L->arena, double, (size_t)n);
    PHF_TRY(m4kernel_device(L, dx, n, c, dw));
    PHF_TRY(download(L, dw, n, w));
    PHF_HIP(hipStreamSynchronize(L->stream));
    return PHF_OK;
}

static int regrid_entry(int kind, const double *x, const double *y, const double *w, int64_t n,
                        double sigma, double overlap, double x0, double y0, double c,
                        phf_blobset **out);

int phf_regrid_m4(const double *x, const double *y, const double *w, int64_t n, double sigma,
                  double overlap, double x0, double y0, double c, phf_blobset **out)
{
    clear_error();
    return regrid_entry(0, x, y, w, n, sigma, overlap, x0, y0, c, out);
}

int phf_diffusion_tutty(const double *x, const double *y, const double *w, int64_t n,
                        double sigma, double overlap, double x0, double y0, double dt, double nu,
                        phf_blobset **out)
{
    clear_error();
    PHF_REQUIRE(std::isfinite(dt) && std::isfinite(nu) && dt * nu >= 0.0,
                "dt*nu must be finite and >= 0");
    return regrid_entry(1, x, y, w, n, sigma, overlap, x0, y0, std::sqrt(dt * nu), out);
}

static int regrid_entry(int kind, const double *x, const double *y, const double *w, int64_t n,
                        double sigma, double overlap, double x0, double y0, double c,
                        phf_blobset **out)
{
    PHF_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    PHF_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "n out of range (%lld)", (long long)n);
    PHF_REQUIRE(std::isfinite(sigma) && sigma > 0 && std::isfinite(overlap) && overlap > 0,
                "sigma and overlap must be finite and > 0");
    PHF_REQUIRE(std::isfinite(x0) && std::isfinite(y0) && std::isfinite(c),
                "bounds and c must be finite");
    PHF_TRY(check_host_array(x, n, "xBlob"));
    PHF_TRY(check_host_array(y, n, "yBlob"));
    PHF_TRY(check_host_array(w, n, "wBlob"));
    phf_blobset_s *res = new phf_blobset_s();
    if (n == 0) {
        *out = res;
        return PHF_OK;
    }
    LaneGuard g;
    if (g.rc != PHF_OK) {
        delete res;
        return g.rc;
    }
    Lane *L = g.lane;
    int rc = [&]() -> int {
        PHF_TRY(reset_arena(L));
        StageTimer tm(L, timing_enabled());
        bool used[PHF_STAGE_COUNT] = {};
        used[PHF_STAGE_REMESH] = used[PHF_STAGE_TOTAL] = true;
        tm.begin(PHF_STAGE_TOTAL);
        double *dx, *dy, *dw;
        PHF_TRY(upload(L, x, n, &dx));
        PHF_TRY(upload(L, y, n, &dy));
        PHF_TRY(upload(L, w, n, &dw));
        tm.begin(PHF_STAGE_REMESH);
        RemeshOut ro;
        ro.want_indices = true;
        PHF_TRY(remesh_device(L, dx, dy, dw, n, sigma, overlap, x0, y0, c, &ro, kind));
        tm.end(PHF_STAGE_REMESH);
        res->m = ro.m;
        res->x.resize(ro.m);
        res->y.resize(ro.m);
        res->w.resize(ro.m);
        res->i.resize(ro.m);
        res->j.resize(ro.m);
        if (ro.m > 0) {
            PHF_TRY(download(L, ro.x, ro.m, res->x.data()));
            PHF_TRY(download(L, ro.y, ro.m, res->y.data()));
            PHF_TRY(download(L, ro.w, ro.m, res->w.data()));
            PHF_HIP(hipMemcpyAsync(res->i.data(), ro.i, sizeof(int64_t) * ro.m,
                                   hipMemcpyDeviceToHost, L->stream));
            PHF_HIP(hipMemcpyAsync(res->j.data(), ro.j, sizeof(int64_t) * ro.m,
                                   hipMemcpyDeviceToHost, L->stream));
        }
        tm.end(PHF_STAGE_TOTAL);
        PHF_HIP(hipStreamSynchronize(L->stream));
        record_stage_times(L, used);
        return PHF_OK;
    }();
    if (rc != PHF_OK) {
        delete res;
        return rc;
    }
    *out = res;
    return PHF_OK;
}

int64_t phf_blobset_size(const phf_blobset *b) { return b ? b->m : -1; }

int phf_blobset_copy(const phf_blobset *b, double *x, double *y, double *w, int64_t *i,
                     int64_t *j)
{
    clear_error();
    PHF_REQUIRE(b != nullptr, "blob set is NULL");
    if (b->m == 0) return PHF_OK;
    if (x) std::memcpy(x, b->x.data(), sizeof(double) * b->m);
    if (y) std::memcpy(y, b->y.data(), sizeof(double) * b->m);
    if (w) std::memcpy(w, b->w.data(), sizeof(double) * b->m);
    if (i && !b->i.empty()) std::memcpy(i, b->i.data(), sizeof(int64_t) * b->m);
    if (j && !b->j.empty()) std::memcpy(j, b->j.data(), sizeof(int64_t) * b->m);
    return PHF_OK;
}

void phf_blobset_free(phf_blobset *b) { delete b; }

int phf_population_control(const double *x, const double *y, const double *w, int64_t n,
                           double wThreshold, double wTotalThreshold, phf_blobset **out)
{
    clear_error();
    PHF_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    PHF_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "n out of range (%lld)", (long long)n);
    PHF_REQUIRE(std::isfinite(wThreshold) && wThreshold >= 0 && std::isfinite(wTotalThreshold),
                "thresholds must be finite (wThreshold >= 0)");
    PHF_TRY(check_host_array(x, n, "xBlob"));
    PHF_TRY(check_host_array(y, n, "yBlob"));
    PHF_TRY(check_host_array(w, n, "wBlob", false));
    phf_blobset_s *res = new phf_blobset_s();
    if (n == 0) {
        *out = res;
        return PHF_OK;
    }
    LaneGuard g;
    if (g.rc != PHF_OK) {
        delete res;
        return g.rc;
    }
    Lane *L = g.lane;
    int rc = [&]() -> int {
        PHF_TRY(reset_arena(L));
        double *dx, *dy, *dw, *xo, *yo, *wo;
        PHF_TRY(upload(L, x, n, &dx));
        PHF_TRY(upload(L, y, n, &dy));
        PHF_TRY(upload(L, w, n, &dw));
        PHF_ALLOC(xo, L->arena, double, (size_t)n);
        PHF_ALLOC(yo, L->arena, double, (size_t)n);
        PHF_ALLOC(wo, L->arena, double, (size_t)n);
        int64_t m = 0;
        PHF_TRY(population_control_device(L, dx, dy, dw, n, wThreshold, wTotalThreshold, xo, yo,
                                          wo, &m, nullptr));
        res->m = m;
        res->x.resize(m);
        res->y.resize(m);
        res->w.resize(m);
        if (m > 0) {
            PHF_TRY(download(L, xo, m, res->x.data()));
            PHF_TRY(download(L, yo, m, res->y.data()));
            PHF_TRY(download(L, wo, m, res->w.data()));
        }
        PHF_HIP(hipStreamSynchronize(L->stream));
        return PHF_OK;
    }();
    if (rc != PHF_OK) {
        delete res;
        return rc;
    }
    *out = res;
    return PHF_OK;
}

}  // extern "C"
