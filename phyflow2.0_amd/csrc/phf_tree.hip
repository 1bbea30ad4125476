// phf_tree.hip -- Morton-key quad-tree build on gfx950.
//
// The reference bins blobs into an adaptive median-split tree
// (fmmsort.cpp:441-785, GPU: cudasort.cu:86 partition kernels).  This
// library uses a uniform quad-tree of depth L over a square domain:
//
//   bbox (exact min/max, FP64)  ->  domain x0, y0, side D (rounded up to 5
//   significant bits so box sides and box-centre differences are exact in
//   float32)  ->  Morton keys  ix = clamp(floor((x-x0)*inv_s), 0, 2^L-1)
//   ->  stable LSD radix sort of (key, index), 8 bits per pass  ->  leaf
//   offsets start[k] = lower_bound(sorted keys, k).
//
// Keys, permutation and offsets are integer results and match a CPU
// stable sort bit for bit (the binning uses the same FP64 operations, with
// no contraction possible: a subtract, a multiply, floor).
//
// Radix sort: per-tile digit histograms -> per-digit scan over the tiles
// (one workgroup per digit; the digit totals are scanned again by every
// tile of the scatter) -> stable scatter.  The first pass's histograms come
// from the kernel that forms the keys.  Inside a tile, elements are ranked 256 at a time; a wave finds
// the lanes sharing its digit with 8 ballots ("match any"), and per-wave
// counts in LDS give the cross-wave prefix, so equal keys keep input order.
#include "phf_tree.hpp"
#include "phf_fmm_dev.hpp"

#include <algorithm>
#include <cmath>

namespace phf {

// ----------------------------------------------------------------- bbox --
constexpr int BB_THREADS = 256;

__device__ __forceinline__ double4 bbox_combine(const double4 &a, const double4 &b)
{
    return make_double4(fmin(a.x, b.x), fmax(a.y, b.y), fmin(a.z, b.z), fmax(a.w, b.w));
}

// {xmin, xmax, ymin, ymax} of the np partial boxes, in every thread (all
// NT threads of the workgroup call it; sh holds NT entries)
template <int NT>
__device__ __forceinline__ double4 bbox_reduce_block(const double4 *__restrict__ part, int np,
                                                     double4 *sh)
{
    double4 acc = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);
    for (int i = threadIdx.x; i < np; i += NT) acc = bbox_combine(acc, part[i]);
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] = bbox_combine(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    const double4 r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(BB_THREADS) void bbox_partial(
    const double *__restrict__ x1, const double *__restrict__ y1, int64_t n1,
    const double *__restrict__ x2, const double *__restrict__ y2, int64_t n2,
    double4 *__restrict__ part)
{
    __shared__ double4 sh[BB_THREADS];
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    const int64_t n = n1 + n2;
    for (int64_t i = (int64_t)blockIdx.x * BB_THREADS + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * BB_THREADS) {
        double x = i < n1 ? x1[i] : x2[i - n1];
        double y = i < n1 ? y1[i] : y2[i - n1];
        xmin = fmin(xmin, x);
        xmax = fmax(xmax, x);
        ymin = fmin(ymin, y);
        ymax = fmax(ymax, y);
    }
    sh[threadIdx.x] = make_double4(xmin, xmax, ymin, ymax);
    __syncthreads();
    for (int s = BB_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            double4 a = sh[threadIdx.x], b = sh[threadIdx.x + s];
            sh[threadIdx.x] = make_double4(fmin(a.x, b.x), fmax(a.y, b.y), fmin(a.z, b.z),
                                           fmax(a.w, b.w));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(BB_THREADS) void bbox_final(const double4 *__restrict__ part,
                                                         int np, double4 *__restrict__ out)
{
    __shared__ double4 sh[BB_THREADS];
    const double4 r = bbox_reduce_block<BB_THREADS>(part, np, sh);
    if (threadIdx.x == 0) *out = r;
}

int device_bbox_parts(Lane *L, const double *x1, const double *y1, int64_t n1, const double *x2,
                      const double *y2, int64_t n2, double4 **parts, int *np)
{
    const int64_t n = n1 + n2;
    int nblk = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + BB_THREADS * 8 - 1) /
                                                                     (BB_THREADS * 8)));
    double4 *part;
    PHF_ALLOC(part, L->arena, double4, (size_t)nblk);
    bbox_partial<<<nblk, BB_THREADS, 0, L->stream>>>(x1, y1, n1, x2, y2, n2, part);
    PHF_HIP(hipGetLastError());
    *parts = part;
    *np = nblk;
    return PHF_OK;
}

int device_bbox(Lane *L, const double *x1, const double *y1, int64_t n1, const double *x2,
                const double *y2, int64_t n2, double4 *d_out)
{
    double4 *part;
    int nblk;
    PHF_TRY(device_bbox_parts(L, x1, y1, n1, x2, y2, n2, &part, &nblk));
    bbox_final<<<1, BB_THREADS, 0, L->stream>>>(part, nblk, d_out);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

Domain make_domain(double xmin, double xmax, double ymin, double ymax, int L, double side)
{
    Domain d{};
    d.x0 = xmin;
    d.y0 = ymin;
    d.D = side > 0.0 ? side : domain_side(std::max(xmax - xmin, ymax - ymin), 0.0);
    d.L = L;
    d.s = d.D / std::ldexp(1.0, L);  // exact (power-of-two division)
    d.inv_s = std::ldexp(1.0, L) / d.D;
    return d;
}

// ----------------------------------------------------------------- keys --
__device__ __forceinline__ uint32_t morton_key(double x, double y, double x0, double y0,
                                               double inv_s, int L)
{
    const double hi = (double)((1u << L) - 1u);
    double fx = floor((x - x0) * inv_s);
    double fy = floor((y - y0) * inv_s);
    fx = (fx >= 0.0) ? fx : 0.0;
    fy = (fy >= 0.0) ? fy : 0.0;
    fx = fx > hi ? hi : fx;
    fy = fy > hi ? hi : fy;
    return morton2((uint32_t)fx, (uint32_t)fy);
}

// ----------------------------------------------------------------- scan --
// Exclusive scan of uint32 (device-wide, three passes).
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_CHUNK = SC_THREADS * SC_ITEMS;

// exclusive scan over the NT threads of the workgroup (all call it;
// sh_wave holds NT / 64 entries)
template <int NT = SC_THREADS>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *sh_wave,
                                                         uint32_t *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) sh_wave[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (int w = 0; w < NT / 64; ++w) {
        uint32_t s = sh_wave[w];
        if (w < wave) before += s;
        all += s;
    }
    __syncthreads();
    if (total) *total = all;
    return before + inc - v;
}

// With dn set, the scan covers the first *dn entries (a count on the
// device, at most the n the grid was sized for); workgroups past it return.
__global__ __launch_bounds__(SC_THREADS) void scan_reduce(const uint32_t *__restrict__ in,
                                                          int64_t n, uint32_t *__restrict__ sums,
                                                          const int64_t *__restrict__ dn)
{
    __shared__ uint32_t shw[SC_THREADS / 64];
    if (dn) n = *dn;
    if ((int64_t)blockIdx.x * SC_CHUNK >= n && blockIdx.x > 0) return;
    const int64_t base = (int64_t)blockIdx.x * SC_CHUNK + (int64_t)threadIdx.x * SC_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t tot;
    block_exclusive_scan(s, shw, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SC_THREADS) void scan_top(uint32_t *__restrict__ sums, int64_t nb,
                                                       const int64_t *__restrict__ dn)
{
    __shared__ uint32_t shw[SC_THREADS / 64];
    if (dn) {
        nb = (*dn + SC_CHUNK - 1) / SC_CHUNK;
        nb = nb > 0 ? nb : 1;
    }
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += SC_THREADS) {
        int64_t i = b0 + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(v, shw, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(SC_THREADS) void scan_apply(const uint32_t *__restrict__ in,
                                                         int64_t n,
                                                         const uint32_t *__restrict__ sums,
                                                         uint32_t *__restrict__ out,
                                                         const int64_t *__restrict__ dn)
{
    __shared__ uint32_t shw[SC_THREADS / 64];
    if (dn) n = *dn;
    if ((int64_t)blockIdx.x * SC_CHUNK >= n) return;
    const int64_t base = (int64_t)blockIdx.x * SC_CHUNK + (int64_t)threadIdx.x * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    uint32_t ex = block_exclusive_scan(s, shw, nullptr) + sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SC_ITEMS; ++k) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
}

int64_t scan_sums_size(int64_t n) { return (n + SC_CHUNK - 1) / SC_CHUNK + 1; }

int device_exclusive_scan(Lane *L, const uint32_t *in, int64_t n, uint32_t *out)
{
    if (n <= 0) return PHF_OK;
    uint32_t *sums;
    PHF_ALLOC(sums, L->arena, uint32_t, (size_t)scan_sums_size(n));
    return device_exclusive_scan_ws(L, in, n, out, sums);
}

int device_exclusive_scan_ws(Lane *L, const uint32_t *in, int64_t n, uint32_t *out,
                             uint32_t *sums, const int64_t *dn)
{
    if (n <= 0) return PHF_OK;
    int64_t nb = (n + SC_CHUNK - 1) / SC_CHUNK;
    scan_reduce<<<(unsigned)nb, SC_THREADS, 0, L->stream>>>(in, n, sums, dn);
    scan_top<<<1, SC_THREADS, 0, L->stream>>>(sums, nb, dn);
    scan_apply<<<(unsigned)nb, SC_THREADS, 0, L->stream>>>(in, n, sums, out, dn);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

// ---------------------------------------------------------- radix sort --
// One workgroup of RS_THREADS (16 waves) per tile of RS_TILE keys, taken in
// RS_ROUNDS rounds of RS_THREADS consecutive keys.
constexpr int RS_THREADS = 1024;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ROUNDS = 4;
constexpr int RS_TILE = RS_THREADS * RS_ROUNDS;
constexpr int RS_SCAN = 256;  // workgroup of rs_rowscan

// Corner of the Morton grid: the device bounding box, or, with parts set,
// the reduction of the np partial boxes of bbox_partial, which workgroup 0
// also stores at *org for the kernels that follow (and at *org_host if
// set).  All threads call it.
__device__ __forceinline__ void tile_origin(const double4 *org, const double4 *__restrict__ parts,
                                            int np, double4 *org_host, double &x0, double &y0)
{
    if (!parts) {
        x0 = org->x;
        y0 = org->z;
        return;
    }
    __shared__ double4 sh[RS_THREADS];
    const double4 bb = bbox_reduce_block<RS_THREADS>(parts, np, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *const_cast<double4 *>(org) = bb;
        if (org_host) *org_host = bb;
    }
    x0 = bb.x;
    y0 = bb.z;
}

// Morton keys of the workgroup's tile of RS_TILE points (L + fine bits per
// axis, inv_sf = 2^fine inv_s; the leaf is key >> 2 fine) and either the
// input index (pay == nullptr) or the float32 record (x - cx, y - cy, w,
// input index bits) relative to the point's leaf centre (cx, cy); with
// hist set, the tile's histogram of the first radix digit (key & mask) in
// the layout of rs_hist.
__global__ __launch_bounds__(RS_THREADS) void morton_tiles(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ w,
    int64_t n, const double4 *org, const double4 *__restrict__ parts, int np,
    double4 *org_host, double s, double inv_sf, int L, int fine, uint32_t *__restrict__ keys,
    int32_t *__restrict__ vals, float4 *__restrict__ pay, uint32_t *__restrict__ hist,
    int nblocks, uint32_t mask)
{
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x;
    double x0, y0;
    tile_origin(org, parts, np, org_host, x0, y0);
    if (tid < 256) h[tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + (int64_t)r * RS_THREADS + tid;
        if (i < n) {
            const double xi = x[i], yi = y[i];
            const uint32_t key = morton_key(xi, yi, x0, y0, inv_sf, L + fine);
            keys[i] = key;
            if (pay) {
                double cx, cy;
                leaf_center(key >> (2 * fine), x0, y0, s, cx, cy);
                pay[i] = make_float4((float)(xi - cx), (float)(yi - cy), (float)w[i],
                                     __int_as_float((int32_t)i));
            } else {
                vals[i] = (int32_t)i;
            }
            if (hist) atomicAdd(&h[key & mask], 1u);
        }
    }
    __syncthreads();
    if (hist && tid < 256) hist[(int64_t)tid * nblocks + blockIdx.x] = h[tid];
}

// Tile histograms of the digit (key >> shift) & mask: hist[d][tile]
// (digit-major, nblocks tiles per row).
// dbits (nullable, device): passes with shift >= *dbits do nothing here and
// copy their tile in rs_scatter.
__global__ __launch_bounds__(RS_THREADS) void rs_hist(const uint32_t *__restrict__ keys,
                                                      int64_t n, int shift, uint32_t mask,
                                                      uint32_t *__restrict__ hist, int nblocks,
                                                      const int32_t *__restrict__ dbits)
{
    __shared__ uint32_t h[256];
    if (dbits && shift >= *dbits) return;
    const int tid = threadIdx.x;
    if (tid < 256) h[tid] = 0;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    uint32_t kr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + (int64_t)r * RS_THREADS + tid;
        kr[r] = i < n ? keys[i] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + (int64_t)r * RS_THREADS + tid;
        if (i < n) atomicAdd(&h[(kr[r] >> shift) & mask], 1u);
    }
    __syncthreads();
    if (tid < 256) hist[(int64_t)tid * nblocks + blockIdx.x] = h[tid];
}

// Row d of the tile histograms (workgroup d): offs[d][b] = tiles before b,
// tot[d] = all tiles.
__global__ __launch_bounds__(RS_SCAN) void rs_rowscan(const uint32_t *__restrict__ hist,
                                                      int nblocks, uint32_t *__restrict__ offs,
                                                      uint32_t *__restrict__ tot, int shift,
                                                      const int32_t *__restrict__ dbits)
{
    __shared__ uint32_t shw[RS_SCAN / 64];
    if (dbits && shift >= *dbits) return;
    const int64_t row = (int64_t)blockIdx.x * nblocks;
    uint32_t carry = 0;
    for (int b0 = 0; b0 < nblocks; b0 += RS_SCAN) {
        const int b = b0 + (int)threadIdx.x;
        const uint32_t v = b < nblocks ? hist[row + b] : 0u;
        uint32_t t;
        const uint32_t ex = block_exclusive_scan<RS_SCAN>(v, shw, &t);
        if (b < nblocks) offs[row + b] = carry + ex;
        carry += t;
    }
    if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// Stable scatter of the tile: digit d's run starts at the digit base (the
// exclusive scan of tot) plus offs[d][tile].  Per round, each wave ranks
// its keys among the lanes with the same digit (nbits ballots) and counts
// them per digit; one thread per digit turns the 16 waves' counts into
// positions in the tile's own sorted order (its digit starts from
// hist[d][tile]), so equal keys keep their input order.  The tile is
// sorted in LDS first and then written out in that order, consecutive
// threads to consecutive places of each digit's run.
template <class VT>
__global__ __launch_bounds__(RS_THREADS) void rs_scatter(
    const uint32_t *__restrict__ kin, const VT *__restrict__ vin, uint32_t *__restrict__ kout,
    VT *__restrict__ vout, int64_t n, int shift, int nbits, const uint32_t *__restrict__ hist,
    const uint32_t *__restrict__ offs, const uint32_t *__restrict__ tot, int nblocks,
    const int32_t *__restrict__ dbits)
{
    __shared__ uint32_t lrun[256], lstart[256], gbase[256];
    __shared__ uint32_t cnt[RS_WAVES][256];
    __shared__ uint32_t pre[RS_WAVES][256];
    __shared__ uint32_t shw[RS_WAVES];
    __shared__ uint32_t sk[RS_TILE];
    __shared__ VT sv[RS_TILE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t mask = (1u << nbits) - 1u;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    const int tile_n = (int)min((int64_t)RS_TILE, n - base);
    if (dbits && shift >= *dbits) {  // nothing left to sort: keep the order
        for (int j = tid; j < tile_n; j += RS_THREADS) {
            kout[base + j] = kin[base + j];
            vout[base + j] = vin[base + j];
        }
        return;
    }
    uint32_t kr[RS_ROUNDS];
    VT vr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + (int64_t)r * RS_THREADS + tid;
        kr[r] = i < n ? kin[i] : 0u;
        vr[r] = i < n ? vin[i] : VT{};
    }
    const bool digit = tid < (1 << nbits);  // tid < 256
    const int64_t hrow = (int64_t)tid * nblocks + blockIdx.x;
    const uint32_t dbase = block_exclusive_scan<RS_THREADS>(digit ? tot[tid] : 0u, shw, nullptr);
    const uint32_t lbase = block_exclusive_scan<RS_THREADS>(digit ? hist[hrow] : 0u, shw, nullptr);
    if (tid < 256) {
        gbase[tid] = dbase + (digit ? offs[hrow] : 0u);
        lstart[tid] = lrun[tid] = lbase;
    }
#pragma unroll
    for (int k = 0; k < RS_WAVES * 256 / RS_THREADS; ++k) (&cnt[0][0])[tid + k * RS_THREADS] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t i = base + (int64_t)r * RS_THREADS + tid;
        const bool valid = i < n;
        const uint32_t key = kr[r];
        const uint32_t d = (key >> shift) & mask;
        uint64_t m = __ballot(valid);
        for (int b = 0; b < nbits; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bal = __ballot(valid && bit);
            m &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0) cnt[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (tid < 256) {
            // tile positions of each wave's first key of digit tid, in wave order
            uint32_t c[RS_WAVES];
#pragma unroll
            for (int w = 0; w < RS_WAVES; ++w) c[w] = cnt[w][tid];
            uint32_t s = lrun[tid];
#pragma unroll
            for (int w = 0; w < RS_WAVES; ++w) {
                pre[w][tid] = s;
                cnt[w][tid] = 0;
                s += c[w];
            }
            lrun[tid] = s;
        }
        __syncthreads();
        if (valid) {
            const uint32_t lp = pre[wave][d] + rank;
            sk[lp] = key;
            sv[lp] = vr[r];
        }
    }
    __syncthreads();
    for (int j = tid; j < tile_n; j += RS_THREADS) {
        const uint32_t key = sk[j];
        const uint32_t d = (key >> shift) & mask;
        const uint32_t pos = gbase[d] + ((uint32_t)j - lstart[d]);
        kout[pos] = key;
        vout[pos] = sv[j];
    }
}

int64_t radix_ws_size(int64_t n)
{
    const int64_t nblocks = std::max<int64_t>(1, (n + RS_TILE - 1) / RS_TILE);
    return 256 * nblocks + 256;  // tile histograms, then the digit totals
}

static int radix_tiles(int64_t n) { return (int)std::max<int64_t>(1, (n + RS_TILE - 1) / RS_TILE); }

int device_radix_sort(Lane *L, uint32_t *keys, int32_t *vals, uint32_t *ktmp, int32_t *vtmp,
                      int64_t n, int key_bits, uint32_t **kres, int32_t **vres)
{
    *kres = keys;
    *vres = vals;
    if (n <= 1 || key_bits <= 0) return PHF_OK;
    const int64_t ws = radix_ws_size(n);
    uint32_t *hist, *offs;
    PHF_ALLOC(hist, L->arena, uint32_t, (size_t)ws);
    PHF_ALLOC(offs, L->arena, uint32_t, (size_t)ws);
    return device_radix_sort_ws(L, keys, vals, ktmp, vtmp, n, key_bits, hist, offs, kres, vres);
}

// hist0: the first pass's tile histograms are already in hist (morton_tiles)
template <class VT>
static int radix_sort_ws(Lane *L, uint32_t *keys, VT *vals, uint32_t *ktmp, VT *vtmp, int64_t n,
                         int key_bits, uint32_t *hist, uint32_t *offs, uint32_t **kres,
                         VT **vres, bool hist0 = false, const int32_t *dbits = nullptr)
{
    *kres = keys;
    *vres = vals;
    if (n <= 1 || key_bits <= 0) return PHF_OK;
    const int nblocks = radix_tiles(n);
    uint32_t *tot = hist + (int64_t)256 * nblocks;
    uint32_t *ka = keys, *kb = ktmp;
    VT *va = vals, *vb = vtmp;
    for (int shift = 0; shift < key_bits; shift += 8) {
        const int nbits = std::min(8, key_bits - shift);
        const uint32_t mask = (1u << nbits) - 1u;
        if (shift > 0 || !hist0)
            rs_hist<<<nblocks, RS_THREADS, 0, L->stream>>>(ka, n, shift, mask, hist, nblocks,
                                                           dbits);
        rs_rowscan<<<1u << nbits, RS_SCAN, 0, L->stream>>>(hist, nblocks, offs, tot, shift,
                                                           dbits);
        rs_scatter<VT><<<nblocks, RS_THREADS, 0, L->stream>>>(ka, va, kb, vb, n, shift, nbits,
                                                              hist, offs, tot, nblocks, dbits);
        PHF_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    *kres = ka;
    *vres = va;
    return PHF_OK;
}

int device_radix_sort_ws(Lane *L, uint32_t *keys, int32_t *vals, uint32_t *ktmp, int32_t *vtmp,
                         int64_t n, int key_bits, uint32_t *hist, uint32_t *offs,
                         uint32_t **kres, int32_t **vres)
{
    return radix_sort_ws<int32_t>(L, keys, vals, ktmp, vtmp, n, key_bits, hist, offs, kres, vres);
}

int device_radix_sort_dbits(Lane *L, uint32_t *keys, int32_t *vals, uint32_t *ktmp, int32_t *vtmp,
                            int64_t n, int key_bits, const int32_t *dbits, uint32_t **kres,
                            int32_t **vres)
{
    *kres = keys;
    *vres = vals;
    if (n <= 1 || key_bits <= 0) return PHF_OK;
    uint32_t *hist, *offs;
    PHF_ALLOC(hist, L->arena, uint32_t, (size_t)radix_ws_size(n));
    PHF_ALLOC(offs, L->arena, uint32_t, (size_t)radix_ws_size(n));
    return radix_sort_ws<int32_t>(L, keys, vals, ktmp, vtmp, n, key_bits, hist, offs, kres, vres,
                                  false, dbits);
}

// --------------------------------------------------------- leaf offsets --
// start[k] = first position of sorted keys with leaf (key >> fsh) >= k
// (k = 0..nleaf); with split set, also the leaf ranges of the 8 XCDs
// holding equal numbers of points: split[j] (j = 0..8) is the leaf of
// sorted point j ceil(n/8), nleaf when there is none.
__global__ void leaf_starts_kernel(const uint32_t *__restrict__ skeys, int64_t n, int64_t nleaf,
                                   int fsh, int32_t *__restrict__ start,
                                   int32_t *__restrict__ split)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (split && k <= 8) {
        const int64_t T = n > 0 ? (n + 7) / 8 : 1;
        split[k] = (k < 8 && k * T < n) ? (int32_t)(skeys[k * T] >> fsh) : (int32_t)nleaf;
    }
    if (k > nleaf) return;
    int64_t lo = 0, hi = n;  // first index with leaf >= k
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((int64_t)(skeys[mid] >> fsh) < k)
            lo = mid + 1;
        else
            hi = mid;
    }
    start[k] = (int32_t)lo;
}

int leaf_starts_device(Lane *L, const uint32_t *skeys, int64_t n, int64_t nkeys, int32_t *start,
                       int32_t *split, int fine)
{
    leaf_starts_kernel<<<div_up(std::max<int64_t>(nkeys + 1, 9), 256), 256, 0, L->stream>>>(
        skeys, n, nkeys, 2 * fine, start, split);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

// Keys and values (index or float32 record) in sorted order, leaf offsets
// and, if asked for, the XCD leaf ranges.
template <class VT>
static int build_sorted(Lane *L, const Domain &dom, const double *x, const double *y,
                        const double *w, int64_t n, SortedPoints *out)
{
    constexpr bool PAY = sizeof(VT) == sizeof(float4);
    uint32_t *k0, *k1;
    VT *v0, *v1;
    const size_t m = (size_t)(n > 0 ? n : 1);
    PHF_ALLOC(k0, L->arena, uint32_t, m);
    PHF_ALLOC(k1, L->arena, uint32_t, m);
    PHF_ALLOC(v0, L->arena, VT, m);
    PHF_ALLOC(v1, L->arena, VT, m);
    const int64_t nleaf = int64_t(1) << (2 * dom.L);
    int32_t *start;
    PHF_ALLOC(start, L->arena, int32_t, (size_t)nleaf + 1);
    out->split = nullptr;
    if (out->want_split) PHF_ALLOC(out->split, L->arena, int32_t, 9);
    const int key_bits = 2 * (dom.L + dom.fine);
    const bool sort = n > 1 && key_bits > 0;
    uint32_t *hist = nullptr, *offs = nullptr;
    if (sort) {
        const int64_t ws = radix_ws_size(n);
        PHF_ALLOC(hist, L->arena, uint32_t, (size_t)ws);
        PHF_ALLOC(offs, L->arena, uint32_t, (size_t)ws);
    }
    const int nblocks = radix_tiles(n);
    if (n > 0)
        morton_tiles<<<nblocks, RS_THREADS, 0, L->stream>>>(
            x, y, w, n, dom.org, dom.parts, dom.nparts, dom.org_host, dom.s,
            std::ldexp(dom.inv_s, dom.fine), dom.L, dom.fine, k0,
            PAY ? nullptr : (int32_t *)v0, PAY ? (float4 *)v0 : nullptr, hist, nblocks,
            (1u << std::min(8, key_bits)) - 1u);
    PHF_HIP(hipGetLastError());
    out->keys_in = nullptr;
    if (out->want_input_keys) {
        PHF_ALLOC(out->keys_in, L->arena, uint32_t, m);
        if (n > 0)
            PHF_HIP(hipMemcpyAsync(out->keys_in, k0, sizeof(uint32_t) * n,
                                   hipMemcpyDeviceToDevice, L->stream));
    }
    uint32_t *ks = k0;
    VT *vs = v0;
    if (sort) PHF_TRY(radix_sort_ws<VT>(L, k0, v0, k1, v1, n, key_bits, hist, offs, &ks, &vs, true));
    PHF_TRY(leaf_starts_device(L, ks, n, nleaf, start, out->split, dom.fine));
    out->fine = dom.fine;
    out->n = n;
    out->keys = ks;
    out->start = start;
    if (PAY) {
        out->perm = nullptr;
        out->pay = (float4 *)vs;
    } else {
        out->perm = (int32_t *)vs;
        out->pay = nullptr;
    }
    return PHF_OK;
}

int build_sorted_points(Lane *L, const Domain &dom, const double *x, const double *y, int64_t n,
                        SortedPoints *out)
{
    return build_sorted<int32_t>(L, dom, x, y, nullptr, n, out);
}

int build_sorted_payload(Lane *L, const Domain &dom, const double *x, const double *y,
                         const double *w, int64_t n, SortedPoints *out)
{
    return build_sorted<float4>(L, dom, x, y, w, n, out);
}

}  // namespace phf

// ------------------------------------------------------------ C entry ----
using namespace phf;

extern "C" int phf_tree_build(const double *x, const double *y, int64_t n, int32_t L,
                              uint32_t *keys, int32_t *perm, int32_t *start, double *x0,
                              double *y0, double *inv_s)
{
    clear_error();
    PHF_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "n out of range (%lld)", (long long)n);
    PHF_REQUIRE(L >= 0 && L <= 12, "L must be in [0, 12] (got %d)", L);
    PHF_TRY(check_host_array(x, n, "x"));
    PHF_TRY(check_host_array(y, n, "y"));
    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *Ln = g.lane;
    PHF_TRY(reset_arena(Ln));
    double *dx, *dy;
    PHF_TRY(upload(Ln, x, n, &dx));
    PHF_TRY(upload(Ln, y, n, &dy));
    double4 *bb;
    PHF_ALLOC(bb, Ln->arena, double4, 1);
    double4 hb = make_double4(0, 0, 0, 0);
    if (n > 0) {
        PHF_TRY(device_bbox(Ln, dx, dy, n, nullptr, nullptr, 0, bb));
        PHF_HIP(hipMemcpyAsync(&hb, bb, sizeof(double4), hipMemcpyDeviceToHost, Ln->stream));
        PHF_HIP(hipStreamSynchronize(Ln->stream));
    }
    Domain dom = make_domain(hb.x, hb.y, hb.z, hb.w, L);
    dom.org = bb;
    SortedPoints sp;
    sp.want_input_keys = keys != nullptr;
    PHF_TRY(build_sorted_points(Ln, dom, dx, dy, n, &sp));
    if (keys && n > 0)
        PHF_HIP(hipMemcpyAsync(keys, sp.keys_in, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                               Ln->stream));
    if (perm && n > 0)
        PHF_HIP(hipMemcpyAsync(perm, sp.perm, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                               Ln->stream));
    if (start)
        PHF_HIP(hipMemcpyAsync(start, sp.start, sizeof(int32_t) * ((size_t(1) << (2 * L)) + 1),
                               hipMemcpyDeviceToHost, Ln->stream));
    PHF_HIP(hipStreamSynchronize(Ln->stream));
    if (x0) *x0 = dom.x0;
    if (y0) *y0 = dom.y0;
    if (inv_s) *inv_s = dom.inv_s;
    return PHF_OK;
}
