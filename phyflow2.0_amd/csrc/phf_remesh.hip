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
// GPU design (no float atomics, one sort of n keys instead of 16n):
//   1. each blob's base cell (iL, jL) -> bounds of the touched range
//   2. stable radix sort of blobs by base-cell key (row-major in range)
//   3. a cell is touched iff one of its 16 base cells (i-a, j-b),
//      a, b in 0..3, holds a blob; exclusive scan of touched flags gives
//      each output cell its row-major slot
//   4. each touched cell sums its contributions slot-major (s = 0..15),
//      blobs of one base cell in ascending input index: exactly the order
//      of a stable sort of the reference's COO entries (entry e = s*n + b),
//      so the float64 sums are reproducible run to run and match the CPU
//      restatement.
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

__global__ __launch_bounds__(RM_THREADS) void base_bounds_final(const longlong4 *__restrict__ part,
                                                                int np, longlong4 *__restrict__ out)
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
    if (threadIdx.x == 0) *out = sh[0];
}

// key = (iL - imin) * nj + (jL - jmin), the blob's base cell in the range
__global__ void base_keys(const double *__restrict__ x, const double *__restrict__ y, int64_t n,
                          double x0, double y0, double h, long long imin, long long jmin,
                          long long nj, uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const long long i = base_index(x[b], x0, h) - imin, j = base_index(y[b], y0, h) - jmin;
    keys[b] = (uint32_t)(i * nj + j);
    vals[b] = (int32_t)b;
}

// sorted blob data: (x/h, y/h, w)
__global__ void gather_scaled(const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ w, const int32_t *__restrict__ perm,
                              int64_t n, double h, double3 *__restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t b = perm[k];
    out[k] = make_double3(x[b] / h, y[b] / h, w[b]);
}

// Cells of the output range (ni x nj, range indices offset by imin, jmin;
// base-cell range is (ni-3) x (nj-3)).  touched = any of the 16 base cells
// (i-a, j-b) non-empty.
__global__ void touched_flags(const int32_t *__restrict__ cstart, long long bi, long long bj,
                              long long ni, long long nj, uint32_t *__restrict__ flag)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ni * nj) return;
    long long i, j;
    divmod_cells(c, nj, i, j);
    uint32_t t = 0;
    for (int s = 0; s < 16 && !t; ++s) {
        const long long ii = i - (s & 3), jj = j - (s >> 2);
        if (ii < 0 || jj < 0 || ii >= bi || jj >= bj) continue;
        const int64_t k = ii * bj + jj;
        t = cstart[k + 1] > cstart[k];
    }
    flag[c] = t;
}

__global__ void gather_sum(const int32_t *__restrict__ cstart, const double3 *__restrict__ sb,
                           const uint32_t *__restrict__ flag, const uint32_t *__restrict__ slot,
                           long long bi, long long bj, long long ni, long long nj,
                           long long imin, long long jmin, double h, double x0, double y0,
                           double x0h, double y0h, double c, double c2,
                           double *__restrict__ xo, double *__restrict__ yo,
                           double *__restrict__ wo, int64_t *__restrict__ io,
                           int64_t *__restrict__ jo)
{
    const int64_t cidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cidx >= ni * nj || !flag[cidx]) return;
    long long i, j;
    divmod_cells(cidx, nj, i, j);
    const long long I = i + imin, J = j + jmin;  // absolute cell indices
    const double ui = ((double)I + 0.5) + x0h;
    const double uj = ((double)J + 0.5) + y0h;
    double acc = 0.0;
    bool first = true;
    for (int s = 0; s < 16; ++s) {
        const long long ii = i - (s & 3), jj = j - (s >> 2);
        if (ii < 0 || jj < 0 || ii >= bi || jj >= bj) continue;
        const int64_t k = ii * bj + jj;
        for (int32_t q = cstart[k]; q < cstart[k + 1]; ++q) {
            const double3 p = sb[q];
            const double v = (m4p_dev(ui - p.x, c, c2) * m4p_dev(uj - p.y, c, c2)) * p.z;
            acc = first ? v : acc + v;
            first = false;
        }
    }
    const uint32_t o = slot[cidx];
    const double hh = 0.5 * h;
    wo[o] = acc;
    xo[o] = ((double)I * h + hh) + x0;
    yo[o] = ((double)J * h + hh) + y0;
    if (io) io[o] = I;
    if (jo) jo[o] = J;
}

// Tutty's diffusion-remesh (_evolution.py:325-445): same touched cells and
// summation order as gather_sum, weights tutty_weight(a) x tutty_weight(b)
// for slot (a, b) of the contributing blob; sb holds unscaled (x, y, w).
__global__ void gather_sum_tutty(const int32_t *__restrict__ cstart, const double3 *__restrict__ sb,
                                 const uint32_t *__restrict__ flag, const uint32_t *__restrict__ slot,
                                 long long bi, long long bj, long long ni, long long nj,
                                 long long imin, long long jmin, double h, double x0, double y0,
                                 double c2, double *__restrict__ xo, double *__restrict__ yo,
                                 double *__restrict__ wo, int64_t *__restrict__ io,
                                 int64_t *__restrict__ jo)
{
    const int64_t cidx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cidx >= ni * nj || !flag[cidx]) return;
    long long i, j;
    divmod_cells(cidx, nj, i, j);
    const long long I = i + imin, J = j + jmin;
    const double hh = 0.5 * h;
    double acc = 0.0;
    bool first = true;
    for (int s = 0; s < 16; ++s) {
        const int a = s & 3, b = s >> 2;
        const long long ii = i - a, jj = j - b;
        if (ii < 0 || jj < 0 || ii >= bi || jj >= bj) continue;
        const int64_t k = ii * bj + jj;
        if (cstart[k + 1] <= cstart[k]) continue;
        // node right of the base cell (ii + imin, jj + jmin)
        const double xc = ((double)(ii + imin + 1) * h + hh) + x0;
        const double yc = ((double)(jj + jmin + 1) * h + hh) + y0;
        for (int32_t q = cstart[k]; q < cstart[k + 1]; ++q) {
            const double3 p = sb[q];
            const double v = (tutty_weight(a, (p.x - xc) / h, c2) *
                              tutty_weight(b, (p.y - yc) / h, c2)) * p.z;
            acc = first ? v : acc + v;
            first = false;
        }
    }
    const uint32_t o = slot[cidx];
    wo[o] = acc;
    xo[o] = ((double)I * h + hh) + x0;
    yo[o] = ((double)J * h + hh) + y0;
    if (io) io[o] = I;
    if (jo) jo[o] = J;
}

int remesh_device(Lane *L, const double *x, const double *y, const double *w, int64_t n,
                  double sigma, double overlap, double x0, double y0, double c, RemeshOut *out,
                  int kind)
{
    out->m = 0;
    if (n == 0) return PHF_OK;
    const double h = sigma * overlap;
    // 1. bounds of base cells
    const int nblk = (int)std::min<int64_t>(1024, (n + RM_THREADS * 8 - 1) / (RM_THREADS * 8));
    longlong4 *part, *bnd;
    PHF_ALLOC(part, L->arena, longlong4, (size_t)nblk);
    PHF_ALLOC(bnd, L->arena, longlong4, 1);
    base_bounds_partial<<<nblk, RM_THREADS, 0, L->stream>>>(x, y, n, x0, y0, h, part);
    base_bounds_final<<<1, RM_THREADS, 0, L->stream>>>(part, nblk, bnd);
    PHF_HIP(hipGetLastError());
    longlong4 hb;
    PHF_HIP(hipMemcpyAsync(&hb, bnd, sizeof(hb), hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    const long long bi = hb.y - hb.x + 1, bj = hb.w - hb.z + 1;  // base-cell range
    const long long ni = bi + 3, nj = bj + 3;                    // touched-cell range
    if (!(bi > 0 && bj > 0) || (double)ni * (double)nj > 4.0e9 ||
        (double)ni * (double)nj > 64.0 * (double)n + 1e6) {
        set_error("remesh cell range %lld x %lld too large for %lld blobs", ni, nj, (long long)n);
        return PHF_ERR_UNSUPPORTED;
    }
    // 2. sort blobs by base cell
    uint32_t *k0, *k1;
    int32_t *v0, *v1;
    PHF_ALLOC(k0, L->arena, uint32_t, (size_t)n);
    PHF_ALLOC(k1, L->arena, uint32_t, (size_t)n);
    PHF_ALLOC(v0, L->arena, int32_t, (size_t)n);
    PHF_ALLOC(v1, L->arena, int32_t, (size_t)n);
    base_keys<<<div_up(n, 256), 256, 0, L->stream>>>(x, y, n, x0, y0, h, hb.x, hb.z, bj, k0, v0);
    PHF_HIP(hipGetLastError());
    int key_bits = 1;
    while (key_bits < 32 && (1ull << key_bits) < (unsigned long long)(bi * bj)) ++key_bits;
    uint32_t *ks;
    int32_t *vs;
    PHF_TRY(device_radix_sort(L, k0, v0, k1, v1, n, key_bits, &ks, &vs));
    int32_t *cstart;
    PHF_ALLOC(cstart, L->arena, int32_t, (size_t)(bi * bj + 1));
    PHF_TRY(leaf_starts_device(L, ks, n, bi * bj, cstart));
    double3 *sb;
    PHF_ALLOC(sb, L->arena, double3, (size_t)n);
    // M4' works on x/h; Tutty on x itself
    gather_scaled<<<div_up(n, 256), 256, 0, L->stream>>>(x, y, w, vs, n, kind == 1 ? 1.0 : h, sb);
    // 3. touched cells and their output slots
    const int64_t ncell = ni * nj;
    uint32_t *flag, *slot;
    PHF_ALLOC(flag, L->arena, uint32_t, (size_t)ncell);
    PHF_ALLOC(slot, L->arena, uint32_t, (size_t)ncell);
    touched_flags<<<div_up(ncell, 256), 256, 0, L->stream>>>(cstart, bi, bj, ni, nj, flag);
    PHF_HIP(hipGetLastError());
    PHF_TRY(device_exclusive_scan(L, flag, ncell, slot));
    uint32_t tail[2];
    PHF_HIP(hipMemcpyAsync(&tail[0], slot + ncell - 1, 4, hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipMemcpyAsync(&tail[1], flag + ncell - 1, 4, hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    const int64_t m = (int64_t)tail[0] + tail[1];
    // 4. outputs
    double *xo, *yo, *wo;
    int64_t *io = nullptr, *jo = nullptr;
    if (out->dst_x) {
        PHF_REQUIRE(m <= out->dst_cap, "remesh output of %lld blobs exceeds capacity %lld",
                    (long long)m, (long long)out->dst_cap);
        xo = out->dst_x;
        yo = out->dst_y;
        wo = out->dst_w;
    } else {
        PHF_ALLOC(xo, L->arena, double, (size_t)std::max<int64_t>(m, 1));
        PHF_ALLOC(yo, L->arena, double, (size_t)std::max<int64_t>(m, 1));
        PHF_ALLOC(wo, L->arena, double, (size_t)std::max<int64_t>(m, 1));
    }
    if (out->want_indices) {
        PHF_ALLOC(io, L->arena, int64_t, (size_t)std::max<int64_t>(m, 1));
        PHF_ALLOC(jo, L->arena, int64_t, (size_t)std::max<int64_t>(m, 1));
    }
    if (kind == 1) {
        const double r = c / h;  // c carries hNu = sqrt(nu dt)
        gather_sum_tutty<<<div_up(ncell, 256), 256, 0, L->stream>>>(
            cstart, sb, flag, slot, bi, bj, ni, nj, hb.x, hb.z, h, x0, y0, r * r, xo, yo, wo, io,
            jo);
    } else {
        gather_sum<<<div_up(ncell, 256), 256, 0, L->stream>>>(
            cstart, sb, flag, slot, bi, bj, ni, nj, hb.x, hb.z, h, x0, y0, x0 / h, y0 / h, c,
            c * c, xo, yo, wo, io, jo);
    }
    PHF_HIP(hipGetLastError());
    out->m = m;
    out->x = xo;
    out->y = yo;
    out->w = wo;
    out->i = io;
    out->j = jo;
    out->cells = ncell;
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
