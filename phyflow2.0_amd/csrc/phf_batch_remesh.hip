// phf_batch_remesh.hip -- forward-Euler convection and M4' remesh of every
// system of a batch in one launch sequence (capturable: no allocation, no
// host synchronisation), after the batched velocity evaluation.
//
// Per system s the semantics are those of the single-system path:
//   x <- x + dt (vx + vInf_x)                    (_evolution.py:540-541)
//   Regrid onto the lattice h_s = sigma_s * overlap, origin (x0h, y0h) h_s
//   (_regridding.py:118-218; Blobs.redistribute uses -1.5 h,
//   __blobsClass.py:891-897), duplicates summed in the order of a stable
//   sort of the reference's COO entries, zero-weight touched cells kept.
//
// Segment table for the lattice: system s owns the cell keys
// [cbase_s, cbase_s + ccap_s).  Its touched-cell range (bi+3) x (bj+3),
// computed on the device from its base-cell bounds, must fit ccap_s; a
// system that does not fit is flagged (m_s = -1) and the host remeshes it
// with the single-system path after the graph has run.  One stable radix
// sort of all blobs by global base-cell key keeps systems contiguous and,
// within a cell, blobs in ascending index; one scan of the touched flags
// of all systems gives every output cell its packed slot, so system s's
// output is the packed range [off_s, off_s + m_s).
#include "phf_batch.hpp"
#include "phf_m4_dev.hpp"

#include <algorithm>
#include <climits>
#include <cmath>

namespace phf {

#pragma clang fp contract(off)

namespace {

__device__ __forceinline__ int find_seg64(const int64_t *__restrict__ arr, int S, int64_t g)
{
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (arr[mid] <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void brm_convect(const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ vx, const double *__restrict__ vy,
                            int64_t N, double dt, double ux, double uy, double *__restrict__ xn,
                            double *__restrict__ yn)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    xn[i] = x[i] + dt * (vx[i] + ux);
    yn[i] = y[i] + dt * (vy[i] + uy);
}

// per-chunk base-cell bounds {imin, imax, jmin, jmax}
__global__ __launch_bounds__(256) void brm_bounds_part(const BatchChunk *__restrict__ ch,
                                                       const BatchRm *__restrict__ rm,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       longlong4 *__restrict__ part)
{
    __shared__ longlong4 sh[256];
    const BatchChunk c = ch[blockIdx.x];
    const BatchRm r = rm[c.sys];
    long long imin = LLONG_MAX, imax = LLONG_MIN, jmin = LLONG_MAX, jmax = LLONG_MIN;
    for (int64_t b = c.beg + threadIdx.x; b < c.end; b += 256) {
        const long long i = base_index(x[b], r.x0, r.h), j = base_index(y[b], r.y0, r.h);
        imin = i < imin ? i : imin;
        imax = i > imax ? i : imax;
        jmin = j < jmin ? j : jmin;
        jmax = j > jmax ? j : jmax;
    }
    sh[threadIdx.x] = make_longlong4(imin, imax, jmin, jmax);
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            const longlong4 a = sh[threadIdx.x], q = sh[threadIdx.x + k];
            sh[threadIdx.x] = make_longlong4(a.x < q.x ? a.x : q.x, a.y > q.y ? a.y : q.y,
                                             a.z < q.z ? a.z : q.z, a.w > q.w ? a.w : q.w);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// one wave per system: base-cell bounds from the chunk partials -> touched
// range, capacity check
__global__ __launch_bounds__(64) void brm_bounds(const BatchSys *__restrict__ sys,
                                                 const BatchRm *__restrict__ rm,
                                                 const int32_t *__restrict__ cfirst,
                                                 const longlong4 *__restrict__ part,
                                                 BatchRmDyn *__restrict__ dyn)
{
    const int s = blockIdx.x, lane = threadIdx.x;
    long long imin = LLONG_MAX, imax = LLONG_MIN, jmin = LLONG_MAX, jmax = LLONG_MIN;
    for (int k = cfirst[s] + lane; k < cfirst[s + 1]; k += 64) {
        const longlong4 q = part[k];
        imin = q.x < imin ? q.x : imin;
        imax = q.y > imax ? q.y : imax;
        jmin = q.z < jmin ? q.z : jmin;
        jmax = q.w > jmax ? q.w : jmax;
    }
    for (int d = 32; d > 0; d >>= 1) {
        const long long a = __shfl_xor(imin, d), b = __shfl_xor(imax, d);
        const long long c = __shfl_xor(jmin, d), e = __shfl_xor(jmax, d);
        imin = a < imin ? a : imin;
        imax = b > imax ? b : imax;
        jmin = c < jmin ? c : jmin;
        jmax = e > jmax ? e : jmax;
    }
    if (lane == 0) {
        const int64_t n = sys[s].n;
        const BatchRm r = rm[s];
        BatchRmDyn d{};
        if (n == 0) {
            d.ok = 1;  // nothing to remesh
        } else {
            d.imin = imin;
            d.jmin = jmin;
            d.bi = imax - imin + 1;
            d.bj = jmax - jmin + 1;
            // (bi + 3)(bj + 3) <= ccap, checked without overflow
            const double cells = (double)(d.bi + 3) * (double)(d.bj + 3);
            d.ok = (d.bi > 0 && d.bj > 0 && cells <= (double)r.ccap) ? 1 : 0;
        }
        dyn[s] = d;
    }
}

__global__ void brm_zero_ranges(int2 *__restrict__ a, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = make_int2(0, 0);
}

// The blob range of every occupied cell from the sorted keys: the first
// blob of a run of equal keys writes its start, the last its end (an empty
// cell keeps the range [0, 0)).
__global__ void brm_cell_runs(const uint32_t *__restrict__ ks, int64_t N,
                              int2 *__restrict__ crange)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t k = ks[i];
    if (i == 0 || ks[i - 1] != k) crange[k].x = (int32_t)i;
    if (i == N - 1 || ks[i + 1] != k) crange[k].y = (int32_t)(i + 1);
}

// Cell keys of the blobs, and each blob's (x/h, y/h, w) packed in one
// record so that the pass in sorted order reads one 32-byte record per
// blob instead of three scattered values.
__global__ void brm_keys(const int64_t *__restrict__ poff, int S, const BatchRm *__restrict__ rm,
                         const BatchRmDyn *__restrict__ dyn, const double *__restrict__ x,
                         const double *__restrict__ y, const double *__restrict__ w, int64_t N,
                         uint32_t *__restrict__ keys, int32_t *__restrict__ vals,
                         double4 *__restrict__ pk)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= N) return;
    // the wave's first blob's system serves every blob of the wave inside it
    const uint64_t u = (uint64_t)b;
    const int64_t b0 = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(
                                      (int)(uint32_t)(u >> 32))
                                  << 32) |
                                 (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u));
    int s = find_seg64(poff, S, b0);
    if (!(poff[s] <= b && (s == S - 1 || b < poff[s + 1]))) s = find_seg64(poff, S, b);
    const BatchRm r = rm[s];
    const BatchRmDyn d = dyn[s];
    const double xb = x[b], yb = y[b];
    uint32_t key = (uint32_t)r.cbase;  // a flagged system's blobs share one key
    if (d.ok) {
        const long long i = base_index(xb, r.x0, r.h) - d.imin;
        const long long j = base_index(yb, r.y0, r.h) - d.jmin;
        key = (uint32_t)(r.cbase + i * d.bj + j);
    }
    keys[b] = key;
    vals[b] = (int32_t)b;
    pk[b] = make_double4(xb / r.h, yb / r.h, w[b], 0.0);
}

// sorted blob data: (x/h, y/h, w)
__global__ void brm_gather_scaled(const double4 *__restrict__ pk, const int32_t *__restrict__ perm,
                                  int64_t N, double3 *__restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const double4 q = pk[perm[k]];
    out[k] = make_double3(q.x, q.y, q.z);
}

// Workgroups are dispatched to the 8 XCDs in turn (workgroup b to XCD
// b mod 8), each with its own L2.  This maps them so that XCD x takes a
// contiguous run of cell blocks: neighbouring rows of cells, which read the
// same blobs, then meet in the same L2.
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned G)
{
    const unsigned q = G / 8u, r = G % 8u, x = b % 8u, k = b / 8u;
    return (x < r ? x * (q + 1u) : r * (q + 1u) + (x - r) * q) + k;
}

// touched = one of the 16 base cells (i-a, j-b), a, b in 0..3, holds a blob.
// Cell kernels run BRM_BLOCK threads per workgroup; every system's cell range
// is a whole number of workgroups, whose system bsys names.
constexpr int BRM_BLOCK = 256;

__global__ __launch_bounds__(BRM_BLOCK) void brm_touched(const int32_t *__restrict__ bsys,
                                                         const BatchRm *__restrict__ rm,
                                                         const BatchRmDyn *__restrict__ dyn,
                                                         const int2 *__restrict__ crange,
                                                         int64_t C, uint32_t *__restrict__ flag)
{
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t c = (int64_t)blk * BRM_BLOCK + threadIdx.x;
    if (c >= C) return;
    const int s = bsys[blk];
    const BatchRmDyn d = dyn[s];
    const int64_t base = rm[s].cbase, loc = c - base;
    const long long nj = d.bj + 3;
    uint32_t t = 0;
    if (d.ok && loc < (d.bi + 3) * nj) {
        long long i, j;
        divmod_cells(loc, nj, i, j);
        for (int q = 0; q < 16 && !t; ++q) {
            const long long ii = i - (q & 3), jj = j - (q >> 2);
            if (ii < 0 || jj < 0 || ii >= d.bi || jj >= d.bj) continue;
            const int64_t k = base + ii * d.bj + jj;
            const int2 rg = crange[k];
            t = rg.y > rg.x;
        }
    }
    flag[c] = t;
}

// Sums of the touched cells (i, j0 + t), t < len <= 4, of one row.  They
// share their base cells: columns jj = j0 + 3 .. j0 - 3 of base rows
// i .. i - 3 (all 28 blob ranges loaded first), and each blob's x weight
// serves every cell of the row.  Per cell the contributions arrive in the
// order of the one-cell sum -- jj descending, then base row descending,
// then blobs ascending -- so the sums are the same.  acc starts at -0.0,
// which the first contribution replaces exactly (-0.0 + v == v for every
// v); a touched cell has at least one contribution.
template <bool WEE>
__device__ __forceinline__ void brm_gather_row4(
    const BatchRm &r, const BatchRmDyn &d, long long i, long long j0, int len, int64_t c0,
    const int2 *__restrict__ crange, const double3 *__restrict__ sb,
    const uint32_t *__restrict__ flag, const uint32_t *__restrict__ slot, double c, double c2,
    double *__restrict__ xo, double *__restrict__ yo, double *__restrict__ wo)
{
    bool tf[4], any = false;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        tf[t] = t < len && flag[c0 + t] != 0u;
        any |= tf[t];
    }
    if (!any) return;
    int2 rg[7][4];
#pragma unroll
    for (int cc = 0; cc < 7; ++cc) {
        const long long jj = j0 + 3 - cc;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const long long ii = i - a;
            rg[cc][a] = make_int2(0, 0);
            // column jj feeds the cells t = jj - j0 .. jj - j0 + 3
            if (3 - cc >= len || ii < 0 || jj < 0 || ii >= d.bi || jj >= d.bj) continue;
            rg[cc][a] = crange[r.cbase + ii * d.bj + jj];
        }
    }
    const double ui = ((double)(i + d.imin) + 0.5) + r.x0h;
    double uj[4], acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        uj[t] = ((double)(j0 + t + d.jmin) + 0.5) + r.y0h;
        acc[t] = -0.0;
    }
#pragma unroll
    for (int cc = 0; cc < 7; ++cc) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            for (int32_t p = rg[cc][a].x; p < rg[cc][a].y; ++p) {
                const double3 b = sb[p];
                const double fx = WEE ? m4p_dev(ui - b.x, c, c2) : m4p_dev0(ui - b.x);
#pragma unroll
                for (int t = (3 - cc > 0 ? 3 - cc : 0); t <= (6 - cc < 3 ? 6 - cc : 3); ++t) {
                    const double fy = WEE ? m4p_dev(uj[t] - b.y, c, c2) : m4p_dev0(uj[t] - b.y);
                    acc[t] = acc[t] + (fx * fy) * b.z;
                }
            }
        }
    }
    const double hh = 0.5 * r.h;
    const double xc = ((double)(i + d.imin) * r.h + hh) + r.x0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (!tf[t]) continue;
        const uint32_t o = slot[c0 + t];
        wo[o] = acc[t];
        xo[o] = xc;
        yo[o] = ((double)(j0 + t + d.jmin) * r.h + hh) + r.y0;
    }
}


// Four consecutive cells of a system's touched range per thread (BRM_BLOCK
// cells per 64-lane workgroup); with nj >= 4 they span at most two rows.
// WEE: Wee's diffusion term (c != 0) in the weights.
template <bool WEE>
__global__ __launch_bounds__(BRM_BLOCK / 4) void brm_gather_sum(
    const int32_t *__restrict__ bsys, const BatchRm *__restrict__ rm,
    const BatchRmDyn *__restrict__ dyn, const int2 *__restrict__ crange,
    const double3 *__restrict__ sb, const uint32_t *__restrict__ flag,
    const uint32_t *__restrict__ slot, int64_t C, double c, double c2, double *__restrict__ xo,
    double *__restrict__ yo, double *__restrict__ wo)
{
    const unsigned blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t c0 = (int64_t)blk * BRM_BLOCK + 4 * threadIdx.x;
    if (c0 >= C) return;
    const int s = bsys[blk];
    const BatchRm r = rm[s];
    const BatchRmDyn d = dyn[s];
    if (!d.ok) return;
    const int64_t loc = c0 - r.cbase;
    const long long nj = d.bj + 3, ni = d.bi + 3;
    if (loc >= ni * nj) return;
    long long i, j;
    divmod_cells(loc, nj, i, j);
    int len = (int)(nj - j < 4 ? nj - j : 4);
    int64_t cs = c0;
    // one copy of the row code for both segments
#pragma nounroll
    for (int seg = 0; seg < 2; ++seg) {
        brm_gather_row4<WEE>(r, d, i, j, len, cs, crange, sb, flag, slot, c, c2, xo, yo, wo);
        if (cs + len == c0 + 4 || i + 1 >= ni) break;
        cs += len;
        len = (int)(c0 + 4 - cs);
        ++i;
        j = 0;
    }
}

// per system: packed output offset and count (-1: did not fit)
__global__ void brm_counts(const BatchRm *__restrict__ rm, const BatchRmDyn *__restrict__ dyn,
                           int S, const uint32_t *__restrict__ flag,
                           const uint32_t *__restrict__ slot, int64_t C,
                           int64_t *__restrict__ mo)
{
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= S) return;
    const int64_t total = C > 0 ? (int64_t)slot[C - 1] + flag[C - 1] : 0;
    const int64_t b0 = rm[s].cbase, b1 = b0 + rm[s].ccap;
    const int64_t o0 = b0 < C ? (int64_t)slot[b0] : total;
    const int64_t o1 = b1 < C ? (int64_t)slot[b1] : total;
    mo[2 * s] = o0;
    mo[2 * s + 1] = dyn[s].ok ? o1 - o0 : -1;
}

}  // namespace

int64_t batch_cell_capacity(int64_t n, double cell_factor)
{
    if (n == 0) return 0;
    const double f = cell_factor > 0 ? cell_factor : 2.0;
    return (int64_t)std::ceil(f * (double)n + 16.0 * std::sqrt((double)n)) + 256;
}

int batch_remesh_plan(const BatchPlan &P, const double *sigma, const phf_batch_remesh_params &rp,
                      BatchRemeshPlan *R)
{
    R->p = rp;
    R->rm.assign(P.S, BatchRm{});
    int64_t base = 0;
    for (int s = 0; s < P.S; ++s) {
        BatchRm &r = R->rm[s];
        r.h = sigma[s] * rp.overlap;
        PHF_REQUIRE(std::isfinite(r.h) && r.h > 0, "lattice spacing of system %d is not > 0", s);
        r.x0 = rp.x0h * r.h;
        r.y0 = rp.y0h * r.h;
        r.x0h = r.x0 / r.h;
        r.y0h = r.y0 / r.h;
        r.cbase = base;
        // whole workgroups of cells per system
        r.ccap = (batch_cell_capacity(P.sys[s].n, rp.cell_factor) + BRM_BLOCK - 1) / BRM_BLOCK *
                 BRM_BLOCK;
        base += r.ccap;
    }
    R->C = base;
    PHF_REQUIRE(R->C < (int64_t(1) << 31), "remesh cell capacity %lld too large (lower "
                "cell_factor or split the batch)", (long long)R->C);
    R->key_bits = 1;
    while (R->key_bits < 32 && (int64_t(1) << R->key_bits) < R->C + 1) ++R->key_bits;
    R->cb.resize(P.S);
    for (int s = 0; s < P.S; ++s) R->cb[s] = R->rm[s].cbase;
    R->bsys.resize((size_t)(R->C / BRM_BLOCK));
    for (int s = 0; s < P.S; ++s)
        for (int64_t b = R->rm[s].cbase / BRM_BLOCK; b < (R->rm[s].cbase + R->rm[s].ccap) / BRM_BLOCK;
             ++b)
            R->bsys[(size_t)b] = s;
    return PHF_OK;
}

int batch_remesh_alloc(const BatchPlan &P, const BatchRemeshPlan &R, BatchRemeshBuffers *Bf)
{
    Arena &A = Bf->arena;
    const size_t N = (size_t)std::max<int64_t>(P.N, 1), C = (size_t)std::max<int64_t>(R.C, 1);
    PHF_ALLOC(Bf->rm, A, BatchRm, (size_t)P.S);
    PHF_ALLOC(Bf->dyn, A, BatchRmDyn, (size_t)P.S);
    PHF_ALLOC(Bf->cb, A, int64_t, (size_t)P.S);
    PHF_ALLOC(Bf->bsys, A, int32_t, std::max<size_t>(R.bsys.size(), 1));
    PHF_ALLOC(Bf->xn, A, double, N);
    PHF_ALLOC(Bf->yn, A, double, N);
    PHF_ALLOC(Bf->sb, A, double3, N);
    PHF_ALLOC(Bf->crange, A, int2, C);
    PHF_ALLOC(Bf->flag, A, uint32_t, C);
    PHF_ALLOC(Bf->slot, A, uint32_t, C);
    PHF_ALLOC(Bf->sums, A, uint32_t, (size_t)scan_sums_size((int64_t)C));
    PHF_ALLOC(Bf->xo, A, double, C);
    PHF_ALLOC(Bf->yo, A, double, C);
    PHF_ALLOC(Bf->wo, A, double, C);
    PHF_ALLOC(Bf->mo, A, int64_t, (size_t)2 * P.S);
    PHF_HIP(hipMemcpy(Bf->rm, R.rm.data(), sizeof(BatchRm) * P.S, hipMemcpyHostToDevice));
    PHF_HIP(hipMemcpy(Bf->cb, R.cb.data(), sizeof(int64_t) * P.S, hipMemcpyHostToDevice));
    if (!R.bsys.empty())
        PHF_HIP(hipMemcpy(Bf->bsys, R.bsys.data(), sizeof(int32_t) * R.bsys.size(),
                          hipMemcpyHostToDevice));
    return PHF_OK;
}

int batch_remesh_enqueue(Lane *L, const BatchPlan &P, BatchBuffers &Bv, const BatchRemeshPlan &R,
                         BatchRemeshBuffers &Bf, const double *x, const double *y,
                         const double *w, const double *vx, const double *vy)
{
    const int S = P.S;
    const int64_t N = P.N, C = R.C;
    const double *xs = x, *ys = y;
    if (R.p.dt != 0.0) {
        if (N > 0)
            brm_convect<<<div_up(N, 256), 256, 0, L->stream>>>(x, y, vx, vy, N, R.p.dt, R.p.vinf_x,
                                                               R.p.vinf_y, Bf.xn, Bf.yn);
        xs = Bf.xn;
        ys = Bf.yn;
    }
    if (!P.chunks.empty())
        brm_bounds_part<<<(unsigned)P.chunks.size(), 256, 0, L->stream>>>(
            Bv.chunks, Bf.rm, xs, ys, reinterpret_cast<longlong4 *>(Bv.cpart));
    brm_bounds<<<S, 64, 0, L->stream>>>(Bv.sys, Bf.rm, Bv.cfirst,
                                        reinterpret_cast<const longlong4 *>(Bv.cpart), Bf.dyn);
    PHF_HIP(hipGetLastError());
    uint32_t *ks = Bv.k0;
    int32_t *vs = Bv.v0;
    // every cell's blob range: empty first (a fill kernel, not a memset: the
    // sequence is captured in a graph), then the occupied cells' from the
    // runs of the sorted keys
    brm_zero_ranges<<<(unsigned)std::min<int64_t>(div_up(C, 256), 4096), 256, 0, L->stream>>>(
        Bf.crange, C);
    if (N > 0) {
        brm_keys<<<div_up(N, 256), 256, 0, L->stream>>>(Bv.poff, S, Bf.rm, Bf.dyn, xs, ys, w, N,
                                                        Bv.k0, Bv.v0, Bv.pk);
        PHF_HIP(hipGetLastError());
        // the velocity pass is done with the sort buffers and workspaces
        PHF_TRY(device_radix_sort_ws(L, Bv.k0, Bv.v0, Bv.k1, Bv.v1, N, R.key_bits, Bv.hist,
                                     Bv.offs, &ks, &vs));
        brm_cell_runs<<<div_up(N, 256), 256, 0, L->stream>>>(ks, N, Bf.crange);
    }
    PHF_HIP(hipGetLastError());
    if (N > 0)
        brm_gather_scaled<<<div_up(N, 256), 256, 0, L->stream>>>(Bv.pk, vs, N, Bf.sb);
    if (C > 0) {
        brm_touched<<<(unsigned)(C / BRM_BLOCK), BRM_BLOCK, 0, L->stream>>>(
            Bf.bsys, Bf.rm, Bf.dyn, Bf.crange, C, Bf.flag);
        PHF_HIP(hipGetLastError());
        PHF_TRY(device_exclusive_scan_ws(L, Bf.flag, C, Bf.slot, Bf.sums));
        const unsigned G = (unsigned)(C / BRM_BLOCK);
        if (R.p.c != 0.0)
            brm_gather_sum<true><<<G, BRM_BLOCK / 4, 0, L->stream>>>(
                Bf.bsys, Bf.rm, Bf.dyn, Bf.crange, Bf.sb, Bf.flag, Bf.slot, C, R.p.c,
                R.p.c * R.p.c, Bf.xo, Bf.yo, Bf.wo);
        else
            brm_gather_sum<false><<<G, BRM_BLOCK / 4, 0, L->stream>>>(
                Bf.bsys, Bf.rm, Bf.dyn, Bf.crange, Bf.sb, Bf.flag, Bf.slot, C, 0.0, 0.0, Bf.xo,
                Bf.yo, Bf.wo);
    }
    brm_counts<<<div_up(S, 64), 64, 0, L->stream>>>(Bf.rm, Bf.dyn, S, Bf.flag, Bf.slot, C, Bf.mo);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

}  // namespace phf
