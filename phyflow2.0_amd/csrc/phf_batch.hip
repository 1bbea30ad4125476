// phf_batch.hip -- many independent blob systems of uneven sizes evaluated
// together: one launch per FMM phase covers every system through a segment
// table, and the whole evaluation is captured once as a hipGraph and
// replayed (no host synchronisation inside).
//
// There is no batched path in the reference: every Blobs object calls the
// FMM on its own (fmm2dInteract, fmm.cpp:88), allocating, copying and
// freeing device memory each time (cudaeval.cu:154-342).
//
// Segment table
//   points:  packed arrays; system s owns [poff[s], poff[s+1])
//   leaves:  global leaf index = loff[s] + Morton key in system s (depth L_s)
//   boxes:   level-major; level l of system s starts at lvl[l*(S+1)+s]
// Each system gets its own square domain (computed on the device from its
// bbox, exactly as the single-system path does on the host), its own leaf
// level and its own smoothing parameters.
//
// Lists are not materialised: the near predicate of a box offset (dx, dy)
// at level l,
//   near_l = (|dx| <= ws && |dy| <= ws) || gap_x^2 + gap_y^2 < cutoff^2,
//   gap = max(|d| - 1, 0) * s_l,
// is evaluated in float64 by every kernel that needs it, so P2P and M2L
// agree on near/far by construction; the interaction list at level l is
// {tau : near_{l-1}(floor((tau + parity)/2)) && !near_l(tau)}.
#include "phf_batch.hpp"
#include "phf_fmm.hpp"
#include "phf_fmm_dev.hpp"
#include "phf_p2p_dev.hpp"
#include "phf_remesh.hpp"

#include <map>
#include <mutex>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace phf {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kRMax = 6;                // largest near radius (boxes)
constexpr int kRU = 2 * kRMax + 1;      // radius of interaction-list offsets
constexpr int kMaxLevel = 11;

// largest s in [0, S) with arr[s] <= g (arr non-decreasing, arr[0] <= g)
__device__ __forceinline__ int find_seg(const int64_t *__restrict__ arr, int S, int64_t g)
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

// find_seg where the active lanes of a wave mostly share one segment
// (consecutive points, sorted keys, consecutive boxes): the first active
// lane's segment is searched with wave-uniform loads and serves every lane
// that lies inside it; the others search alone.
__device__ __forceinline__ int find_seg_wave(const int64_t *__restrict__ arr, int S, int64_t g)
{
    const uint64_t u = (uint64_t)g;
    const int64_t g0 = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(
                                      (int)(uint32_t)(u >> 32))
                                  << 32) |
                                 (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u));
    const int s0 = find_seg(arr, S, g0);
    if (arr[s0] <= g && (s0 == S - 1 || g < arr[s0 + 1])) return s0;
    return find_seg(arr, S, g);
}

__device__ __forceinline__ bool near_pred(int dx, int dy, int ws, double s, double c2)
{
    return near_offset(dx, dy, ws, s, c2);
}

__device__ __forceinline__ int near_radius(int ws, double s, double c2)
{
    // smallest R such that every near offset has max(|dx|,|dy|) <= R: an
    // offset with max(|dx|,|dy|) = a > ws is near only through a gap
    // (a - 1) s below the cutoff, so R = ceil(cutoff / s) suffices once the
    // predicate's own float64 test rejects the gap R s (monotone rounding
    // then rejects every longer gap)
#pragma clang fp contract(off)
    int R = (int)ceil(sqrt(c2) / s);
    R = R > 1 ? R : 1;
    while (((double)R * s) * ((double)R * s) < c2) ++R;
    return R > ws ? R : ws;
}

__device__ __forceinline__ int floordiv2(int v) { return v >= 0 ? v / 2 : -((1 - v) / 2); }

// ------------------------------------------------------- domains, keys ----
// Per-system domain: bounding box, side from choose_domain_side (the rule
// the single-system FMM applies, so a system builds the same tree alone and
// in the batch).
// per-chunk bounding boxes {xmin, xmax, ymin, ymax}
__global__ __launch_bounds__(256) void batch_bbox_part(const BatchChunk *__restrict__ ch,
                                                       const double *__restrict__ x,
                                                       const double *__restrict__ y,
                                                       double4 *__restrict__ part)
{
    __shared__ double4 sh[256];
    const BatchChunk c = ch[blockIdx.x];
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (int64_t i = c.beg + threadIdx.x; i < c.end; i += 256) {
        const double a = x[i], b = y[i];
        xmin = fmin(xmin, a);
        xmax = fmax(xmax, a);
        ymin = fmin(ymin, b);
        ymax = fmax(ymax, b);
    }
    sh[threadIdx.x] = make_double4(xmin, xmax, ymin, ymax);
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) {
            const double4 p = sh[threadIdx.x], q = sh[threadIdx.x + k];
            sh[threadIdx.x] = make_double4(fmin(p.x, q.x), fmax(p.y, q.y), fmin(p.z, q.z),
                                           fmax(p.w, q.w));
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Per-system domain: bounding box from the chunk partials, side from
// choose_domain_side (the rule the single-system FMM applies, so a system
// builds the same tree alone and in the batch).
__global__ __launch_bounds__(64) void batch_domain(const BatchSys *__restrict__ sys,
                                                   const int32_t *__restrict__ cfirst,
                                                   const double4 *__restrict__ part, int ws,
                                                   int p1, int chunk, BatchDom *__restrict__ dom)
{
    const int s = blockIdx.x, lane = threadIdx.x;
    double4 b = make_double4(INFINITY, -INFINITY, INFINITY, -INFINITY);
    for (int k = cfirst[s] + lane; k < cfirst[s + 1]; k += 64) {
        const double4 q = part[k];
        b = make_double4(fmin(b.x, q.x), fmax(b.y, q.y), fmin(b.z, q.z), fmax(b.w, q.w));
    }
    for (int d = 32; d > 0; d >>= 1) {
        const double4 q = make_double4(__shfl_xor(b.x, d), __shfl_xor(b.y, d),
                                       __shfl_xor(b.z, d), __shfl_xor(b.w, d));
        b = make_double4(fmin(b.x, q.x), fmax(b.y, q.y), fmin(b.z, q.z), fmax(b.w, q.w));
    }
    if (lane == 0) {
        const BatchSys sy = sys[s];
        const int64_t n = sy.n;
        BatchDom d;
        if (n == 0) b = make_double4(0.0, 0.0, 0.0, 0.0);
        d.x0 = b.x;
        d.y0 = b.z;
        const int L = sy.L;
        d.D = choose_domain_side(b.y - b.x, b.w - b.z, L, (double)n, (double)n, sy.cutoff, ws,
                                 1.4426950408889634074 * sy.inv_ks, p1, chunk);
        d.s = d.D / ldexp(1.0, L);
        d.inv_s = ldexp(1.0, L) / d.D;
        dom[s] = d;
    }
}

// Leaf keys of the points, and each point's (x, y, w) packed in one record
// so that the pass in sorted order reads one 32-byte record per point
// instead of three scattered values.
__global__ void batch_keys(const BatchSys *__restrict__ sys, const int64_t *__restrict__ poff,
                           int S, const BatchDom *__restrict__ dom, const double *__restrict__ x,
                           const double *__restrict__ y, const double *__restrict__ w, int64_t N,
                           uint32_t *__restrict__ keys, int32_t *__restrict__ vals,
                           double4 *__restrict__ pk)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int s = find_seg_wave(poff, S, i);
    const BatchDom d = dom[s];
    const int L = sys[s].L;
    const double hi = (double)((1u << L) - 1u);
    const double xi = x[i], yi = y[i];
    pk[i] = make_double4(xi, yi, w[i], 0.0);
    double fx = floor((xi - d.x0) * d.inv_s);
    double fy = floor((yi - d.y0) * d.inv_s);
    fx = fx >= 0.0 ? fx : 0.0;
    fy = fy >= 0.0 ? fy : 0.0;
    fx = fx > hi ? hi : fx;
    fy = fy > hi ? hi : fy;
    keys[i] = (uint32_t)(sys[s].leafbase + morton2((uint32_t)fx, (uint32_t)fy));
    vals[i] = (int32_t)i;
}

// float32: offsets from the leaf centre into the padded per-leaf ranges
// (sp); float64: absolute positions in sorted order
template <class T>
__global__ void batch_gather(const BatchSys *__restrict__ sys, const int64_t *__restrict__ loff,
                             int S, const BatchDom *__restrict__ dom,
                             const double4 *__restrict__ pk, const uint32_t *__restrict__ skeys,
                             const int32_t *__restrict__ perm, const int32_t *__restrict__ start,
                             const int32_t *__restrict__ sp, int64_t N, void *__restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N) return;
    const double4 q = pk[perm[k]];
    if (sizeof(T) == 4) {
        const int64_t g = skeys[k];
        const int s = find_seg_wave(loff, S, g);
        const BatchDom d = dom[s];
        double cx, cy;
        leaf_center((uint32_t)(g - sys[s].leafbase), d.x0, d.y0, d.s, cx, cy);
        reinterpret_cast<float4 *>(out)[sp[g] + (k - start[g])] =
            make_float4((float)(q.x - cx), (float)(q.y - cy), (float)q.z, 0.0f);
    } else {
        reinterpret_cast<double4 *>(out)[k] = make_double4(q.x, q.y, q.z, 0.0);
    }
}

__global__ void batch_clear_flag(int32_t *flag) { flag[0] = 0; }

// --------------------------------------------------------------- P2M ------
template <class T>
__global__ __launch_bounds__(64) void batch_p2m(const BatchSys *__restrict__ sys,
                                                const int64_t *__restrict__ loff,
                                                const int64_t *__restrict__ lvl, int S,
                                                const BatchDom *__restrict__ dom,
                                                const void *__restrict__ src_,
                                                const int32_t *__restrict__ start, int p1,
                                                typename C2<T>::type *__restrict__ M)
{
    using V = typename C2<T>::type;
    extern __shared__ unsigned char lds_raw[];
    V *tab = reinterpret_cast<V *>(lds_raw);
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    const int s = find_seg(loff, S, g);
    const int L = sys[s].L;
    if (L < 2) return;  // no far field in this system
    const uint32_t b = (uint32_t)(g - sys[s].leafbase);
    const int32_t s0 = start[g], s1 = start[g + 1];
    V acc = cmk<T>(0, 0);
    if (s1 > s0) {
        const BatchDom d = dom[s];
        double cx, cy;
        leaf_center(b, d.x0, d.y0, d.s, cx, cy);
        const T inv_s = (T)(1.0 / d.s);
        for (int32_t c0 = s0; c0 < s1; c0 += 64) {
            const int32_t j = c0 + lane;
            V z = cmk<T>(0, 0), q = cmk<T>(0, 0);
            if (j < s1) {
                if (sizeof(T) == 4) {
                    const float4 p = reinterpret_cast<const float4 *>(src_)[j];
                    z = cmk<T>((T)p.x * inv_s, (T)p.y * inv_s);
                    q = cmk<T>((T)p.z, 0);
                } else {
                    const double4 p = reinterpret_cast<const double4 *>(src_)[j];
                    z = cmk<T>((T)((p.x - cx) * (double)inv_s), (T)((p.y - cy) * (double)inv_s));
                    q = cmk<T>((T)p.z, 0);
                }
            }
            for (int k = 0; k < p1; ++k) {
                tab[lane * p1 + k] = q;
                q = cmul(q, z);
            }
            __syncthreads();
            // term k summed by 64 / W lanes (k, k + W, ...), each over every
            // (64 / W)-th point, then combined by shuffles: a chain of
            // cnt W / 64 adds per lane instead of cnt
            const int cnt = min(64, s1 - c0);
            const int W = p1 <= 16 ? 16 : (p1 <= 32 ? 32 : 64);
            const int k = lane & (W - 1), R = 64 / W;
            V part = cmk<T>(0, 0);
            if (k < p1)
                for (int jj = lane / W; jj < cnt; jj += R) {
                    const V t = tab[jj * p1 + k];
                    part.x += t.x;
                    part.y += t.y;
                }
            for (int o = W; o < 64; o <<= 1) {
                part.x += __shfl_xor(part.x, o);
                part.y += __shfl_xor(part.y, o);
            }
            if (lane < p1) {
                acc.x += part.x;
                acc.y += part.y;
            }
            __syncthreads();
        }
    }
    if (lane < p1) M[(lvl[L * (S + 1) + s] + b) * p1 + lane] = acc;
}

// --------------------------------------------------------- M2M / L2L ------
template <class T, class TC>
__global__ void batch_m2m(const BatchSys *__restrict__ sys, const int64_t *__restrict__ lvl,
                          int S, int l, const int32_t *__restrict__ start, int p1,
                          typename C2<T>::type *__restrict__ M,
                          const typename C2<TC>::type *__restrict__ tab)
{
    using V = typename C2<T>::type;
    using W = typename C2<TC>::type;
    const int64_t *row = lvl + (int64_t)l * (S + 1);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = row[S] - row[0];
    if (t >= nb * p1) return;
    const int64_t gb = row[0] + t / p1;
    const int m = (int)(t % p1);
    const int s = find_seg_wave(row, S, gb);
    const int L = sys[s].L;
    if (l >= L) return;  // level l is not a parent level of system s
    const uint32_t b = (uint32_t)(gb - row[s]);
    W acc = W{0, 0};
    if (box_count(start + sys[s].leafbase, b, 2 * (L - l)) > 0) {
        const int64_t cbase = lvl[(int64_t)(l + 1) * (S + 1) + s];
        for (int c = 0; c < 4; ++c) {
            const V *ch = M + (cbase + 4 * (int64_t)b + c) * p1;
            const W *tr = tab + ((int64_t)c * p1 + m) * p1;
            for (int k = 0; k <= m; ++k) {
                const V a = ch[k];
                const W tk = tr[k];
                acc.x += tk.x * (TC)a.x - tk.y * (TC)a.y;
                acc.y += tk.x * (TC)a.y + tk.y * (TC)a.x;
            }
        }
    }
    M[gb * p1 + m] = V{(T)acc.x, (T)acc.y};
}

template <class T, class TC>
__global__ void batch_l2l(const BatchSys *__restrict__ sys, const int64_t *__restrict__ lvl,
                          int S, int l, const int32_t *__restrict__ start, int p1,
                          typename C2<T>::type *__restrict__ B,
                          const typename C2<TC>::type *__restrict__ tab)
{
    // children at level l+1 receive their parent's local expansion
    using V = typename C2<T>::type;
    using W = typename C2<TC>::type;
    const int64_t *row = lvl + (int64_t)(l + 1) * (S + 1);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nb = row[S] - row[0];
    if (t >= nb * p1) return;
    const int64_t gb = row[0] + t / p1;
    const int m = (int)(t % p1);
    const int s = find_seg_wave(row, S, gb);
    const int L = sys[s].L;
    if (l < 2 || l + 1 > L) return;
    const uint32_t b = (uint32_t)(gb - row[s]);
    if (box_count(start + sys[s].leafbase, b, 2 * (L - l - 1)) == 0) return;
    const V *par = B + (lvl[(int64_t)l * (S + 1) + s] + (b >> 2)) * p1;
    const W *tr = tab + ((int64_t)(b & 3) * p1 + m) * p1;
    W acc = W{0, 0};
    for (int j = m; j < p1; ++j) {
        const V a = par[j];
        const W tk = tr[j];
        acc.x += tk.x * (TC)a.x - tk.y * (TC)a.y;
        acc.y += tk.x * (TC)a.y + tk.y * (TC)a.x;
    }
    const V cur = B[gb * p1 + m];
    B[gb * p1 + m] = V{(T)((TC)cur.x + acc.x), (T)((TC)cur.y + acc.y)};
}

// --------------------------------------------------------------- M2L ------
template <class T, int P1>
__global__ __launch_bounds__(64) void batch_m2l(
    const BatchSys *__restrict__ sys, const int64_t *__restrict__ lvl, int S, int l,
    const BatchDom *__restrict__ dom, const int32_t *__restrict__ start, int p1, int ws,
    const typename C2<T>::type *__restrict__ M, typename C2<T>::type *__restrict__ B,
    const typename C2<T>::type *__restrict__ U, const T *__restrict__ binom,
    int32_t *__restrict__ flag)
{
    using V = typename C2<T>::type;
    __shared__ V Ap[64];
    const int lane = threadIdx.x;
    const int G = 64 / p1;
    const int g = lane / p1, m = lane - g * p1;
    const int64_t *row = lvl + (int64_t)l * (S + 1);
    const int64_t nb = row[S] - row[0];
    const int64_t slot = (int64_t)blockIdx.x * G + g;
    bool active = g < G && slot < nb;
    int s = 0, L = 0, bx = 0, by = 0, par = 0, n = 1 << l;
    uint32_t b = 0;
    double sl = 1.0, slp = 2.0, c2 = 0.0;
    int Rp = 0, Rw = 0;
    int64_t gb = 0;
    if (active) {
        gb = row[0] + slot;
        s = find_seg(row, S, gb);
        L = sys[s].L;
        active = l <= L;
        b = (uint32_t)(gb - row[s]);
        bx = (int)compact_bits(b);
        by = (int)compact_bits(b >> 1);
        par = (bx & 1) | ((by & 1) << 1);
        const double D = dom[s].D;
        sl = D / ldexp(1.0, l);
        slp = 2.0 * sl;
        c2 = sys[s].c2;
        Rp = near_radius(ws, slp, c2);
        if (Rp > kRMax) {
            if (m == 0) atomicOr(flag, 1);
            Rp = kRMax;
        }
        Rw = 2 * Rp + 1;
        if (active) active = box_count(start + sys[s].leafbase, b, 2 * (L - l)) > 0;
    }
    // window of candidate offsets, wave-uniform bound
    int side = active ? 2 * Rw + 1 : 0;
    int ncand = side * side;
    for (int o = 32; o > 0; o >>= 1) ncand = max(ncand, __shfl_xor(ncand, o, 64));
    T crow[P1];
#pragma unroll
    for (int k = 0; k < P1; ++k) crow[k] = (k < p1 && m < p1) ? binom[m * p1 + k] : (T)0;
    const int W2 = 2 * kRU + 1;
    const T sgn = (m & 1) ? (T)-1 : (T)1;
    T ar = 0, ai = 0;
    for (int t = 0; t < ncand; ++t) {
        bool ok = active && t < side * side;
        int tx = 0, ty = 0;
        uint32_t sb = 0;
        if (ok) {
            tx = t % side - Rw;
            ty = t / side - Rw;
            const int sx = bx + tx, sy = by + ty;
            ok = sx >= 0 && sy >= 0 && sx < n && sy < n;
            if (ok) {
                const int px = floordiv2(tx + (par & 1)), py = floordiv2(ty + (par >> 1));
                ok = near_pred(px, py, ws, slp, c2) && !near_pred(tx, ty, ws, sl, c2);
            }
            if (ok) {
                sb = morton2((uint32_t)sx, (uint32_t)sy);
                ok = box_count(start + sys[s].leafbase, sb, 2 * (L - l)) > 0;
            }
        }
        const V *Uo = U + (int64_t)((ty + kRU) * W2 + (tx + kRU)) * (p1 + 1);
        if (ok) Ap[lane] = cmul(M[(row[s] + sb) * p1 + m], Uo[m]);
        __syncthreads();
        if (ok) {
            T sr = 0, si = 0;
            const V *ap = Ap + g * p1;
#pragma unroll
            for (int k = 0; k < P1; ++k)
                if (k < p1) {
                    const V v = ap[k];
                    sr += crow[k] * v.x;
                    si += crow[k] * v.y;
                }
            const V um = Uo[m + 1];
            ar += sgn * (um.x * sr - um.y * si);
            ai += sgn * (um.x * si + um.y * sr);
        }
        __syncthreads();
    }
    if (g < G && slot < nb && m < p1 && l <= L) {
        const T inv_r = (T)(1.0 / sl);
        B[gb * p1 + m] = active ? V{ar * inv_r, ai * inv_r} : V{0, 0};
    }
}

// M2L with the binomial products on the matrix cores (float32, p + 1 <= 17).
// One wave per target box.  The candidate offsets of the window are tested
// 64 at a time, one per lane, and the interaction-list links compacted in
// LDS; for each batch of up to 32 links the scaled multipoles u^k A_k form
// the columns (real and imaginary parts apart) of a 17 x 64 matrix, and
// C = Pascal x columns runs as 16 x 16 x 4 float32 MFMAs for rows and
// terms 0..15, with term and row 16 on the vector ALU.  Each lane then
// multiplies its entries of C by (-1)^m u^{m+1} and the 16 lanes of a row
// sum their link columns.  Same links and operator as batch_m2l; the sum
// over links runs in another order.
constexpr int BM_LINKS = 32;

constexpr int BM_ROWS = 17;  // p + 1 <= 17

__global__ __launch_bounds__(64) void batch_m2l_mfma(
    const BatchSys *__restrict__ sys, const int64_t *__restrict__ lvl, int S, int l,
    const BatchDom *__restrict__ dom, const int32_t *__restrict__ start, int p1, int ws,
    const float2 *__restrict__ M, float2 *__restrict__ B, const float2 *__restrict__ U,
    const float *__restrict__ binom, int32_t *__restrict__ flag)
{
    __shared__ float Bs[BM_ROWS][2 * BM_LINKS + 4];
    __shared__ float2 Us[BM_ROWS][BM_LINKS];  // u^{m+1} of each link, m = 0..p
    __shared__ int32_t lsrc[64];
    __shared__ int32_t loff[64];
    const int lane = threadIdx.x;
    const int64_t *row = lvl + (int64_t)l * (S + 1);
    const int64_t nb = row[S] - row[0];
    const int64_t slot = blockIdx.x;
    if (slot >= nb) return;
    const int64_t gb = row[0] + slot;
    const int s = find_seg_wave(row, S, gb);
    const int L = sys[s].L;
    if (l > L) return;
    const uint32_t b = (uint32_t)(gb - row[s]);
    const int bx = (int)compact_bits(b), by = (int)compact_bits(b >> 1);
    const int par = (bx & 1) | ((by & 1) << 1);
    const int n = 1 << l;
    const double D = dom[s].D;
    const double sl = D / ldexp(1.0, l), slp = 2.0 * sl;
    const double c2 = sys[s].c2;
    int Rp = near_radius(ws, slp, c2);
    if (Rp > kRMax) {
        if (lane == 0) atomicOr(flag, 1);
        Rp = kRMax;
    }
    const int Rw = 2 * Rp + 1, side = 2 * Rw + 1, ncand = side * side;
    const int W2 = 2 * kRU + 1;
    const bool active = box_count(start + sys[s].leafbase, b, 2 * (L - l)) > 0;
    // operator rows in registers: A operand of the 16 x 16 x 4 products
    // (lane holds P[lane % 16][4 ks + lane / 16]), term 16 of rows 4 (lane / 16) + v
    const int mr = lane & 15, kq = lane >> 4;
    float pa[4], p16[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        const int k = 4 * ks + kq;
        pa[ks] = (mr < p1 && k < p1) ? binom[mr * p1 + k] : 0.0f;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int m = 4 * kq + v;
        p16[v] = (p1 == 17 && m < p1) ? binom[m * p1 + 16] : 0.0f;
    }
    // terms p + 1 .. 16 of the staged matrix stay zero
    for (int i = lane; i < (BM_ROWS - p1) * (2 * BM_LINKS + 4); i += 64)
        Bs[p1 + i / (2 * BM_LINKS + 4)][i % (2 * BM_LINKS + 4)] = 0.0f;
    float pr[4] = {0, 0, 0, 0}, pi[4] = {0, 0, 0, 0};
    float qr = 0.0f, qi = 0.0f;  // row 16
    for (int c0 = 0; active && c0 < ncand; c0 += 64) {
        // links of this chunk of candidate offsets, in offset order
        const int t = c0 + lane;
        bool ok = t < ncand;
        int tx = 0, ty = 0;
        uint32_t sb = 0;
        if (ok) {
            tx = t % side - Rw;
            ty = t / side - Rw;
            const int sx = bx + tx, sy = by + ty;
            ok = sx >= 0 && sy >= 0 && sx < n && sy < n;
            if (ok) {
                const int px = floordiv2(tx + (par & 1)), py = floordiv2(ty + (par >> 1));
                ok = near_pred(px, py, ws, slp, c2) && !near_pred(tx, ty, ws, sl, c2);
            }
            if (ok) {
                sb = morton2((uint32_t)sx, (uint32_t)sy);
                ok = box_count(start + sys[s].leafbase, sb, 2 * (L - l)) > 0;
            }
        }
        const uint64_t mask = __ballot(ok);
        const int K = __popcll(mask);
        if (ok) {
            const int pos = __popcll(mask & ((1ull << lane) - 1ull));
            lsrc[pos] = (int32_t)(row[s] + sb);
            loff[pos] = (ty + kRU) * W2 + (tx + kRU);
        }
        __syncthreads();
        for (int j0 = 0; j0 < K; j0 += BM_LINKS) {
            const int Kb = K - j0 < BM_LINKS ? K - j0 : BM_LINKS;
            // stage u^k A_k (columns 2j, 2j + 1) and u^{m+1} of each link
            for (int i = lane; i < BM_LINKS * p1; i += 64) {
                const int j = i / p1, k = i - j * p1;
                float2 a = make_float2(0.0f, 0.0f), um = make_float2(0.0f, 0.0f);
                if (j < Kb) {
                    const float2 *Uo = U + (int64_t)loff[j0 + j] * (p1 + 1);
                    a = cmul(M[(int64_t)lsrc[j0 + j] * p1 + k], Uo[k]);
                    um = Uo[k + 1];
                }
                Bs[k][2 * j] = a.x;
                Bs[k][2 * j + 1] = a.y;
                Us[k][j] = um;
            }
            __syncthreads();
#pragma unroll
            for (int ct = 0; ct < 2 * BM_LINKS / 16; ++ct) {
                const int c = 16 * ct + mr;  // this lane's column of the tile
                floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < 4; ++ks)
                    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(pa[ks], Bs[4 * ks + kq][c], acc, 0,
                                                               0, 0);
                const int j = c >> 1;
                const bool im = c & 1;
                const float b16 = Bs[16][c];  // zero below order 16
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int m = 4 * kq + v;
                    if (m < p1 && m < 16) {
                        const float d = acc[v] + p16[v] * b16;
                        const float2 u = Us[m][j];
                        const float sg = (m & 1) ? -1.0f : 1.0f;
                        // C[m][j] = d (real column) or i d (imaginary column)
                        pr[v] += im ? -sg * d * u.y : sg * d * u.x;
                        pi[v] += im ? sg * d * u.x : sg * d * u.y;
                    }
                }
            }
            if (p1 == 17) {  // row 16: column `lane`
                float d = 0.0f;
#pragma unroll
                for (int k = 0; k < BM_ROWS; ++k) d += binom[16 * BM_ROWS + k] * Bs[k][lane];
                const float2 u = Us[16][lane >> 1];
                // (-1)^16 = 1
                qr += (lane & 1) ? -d * u.y : d * u.x;
                qi += (lane & 1) ? d * u.x : d * u.y;
            }
            __syncthreads();
        }
    }
    // sum over the 16 link columns of each row group, and over all lanes for row 16
#pragma unroll
    for (int o = 1; o < 16; o <<= 1)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            pr[v] += __shfl_xor(pr[v], o, 64);
            pi[v] += __shfl_xor(pi[v], o, 64);
        }
    if (p1 == 17)
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            qr += __shfl_xor(qr, o, 64);
            qi += __shfl_xor(qi, o, 64);
        }
    const float inv_r = (float)(1.0 / sl);
    if (mr == 0)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = 4 * kq + v;
            if (m < p1 && m < 16)
                B[gb * p1 + m] = active ? make_float2(pr[v] * inv_r, pi[v] * inv_r)
                                        : make_float2(0.0f, 0.0f);
        }
    if (p1 == 17 && lane == 0)
        B[gb * p1 + 16] = active ? make_float2(qr * inv_r, qi * inv_r) : make_float2(0.0f, 0.0f);
}

// ------------------------------------------------------ P2P + L2P ---------
// Near set of a leaf of system s: the offsets of the window |d| <= R that
// pass the near predicate, enumerated from a table of the window |d| <= kRMax
// sorted by box gap, so that the offsets whose pairs may lie inside the
// smoothing core come first; offsets whose box gap g gives
// shape g^2 log2(e) >= 26 (2% margin) skip the exponential (exact, see
// pair_f32_far).
struct BatchNear {
    const int2 *win;
    int R, ws, bx, by, n;
    double sd, c2, shape_log2;
    int64_t leafbase;
    float s_f;
    __device__ __forceinline__ bool operator()(int q, uint32_t &sb, float &ox, float &oy,
                                               int &core) const
    {
        const int2 o = win[q];
        const int dx = o.x, dy = o.y;
        ox = (float)dx * s_f;
        oy = (float)dy * s_f;
        const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy;
        const double gx = (double)(ax > 1 ? ax - 1 : 0) * sd * 0.99;
        const double gy = (double)(ay > 1 ? ay - 1 : 0) * sd * 0.99;
        core = !(shape_log2 * (gx * gx + gy * gy) >= 26.0);
        if (ax > R || ay > R) return false;
        const int sx = bx + dx, sy = by + dy;
        if (sx < 0 || sy < 0 || sx >= n || sy >= n) return false;
        if (!near_pred(dx, dy, ws, sd, c2)) return false;
        sb = (uint32_t)(leafbase + morton2((uint32_t)sx, (uint32_t)sy));
        return true;
    }
};

// P2P + L2P of leaf gl of the whole batch (all threads of the workgroup):
// sources staged in LDS (p2p_leaf); targets are the system's own blobs,
// read from the padded source array.
template <int KIND>
__device__ __forceinline__ void batch_p2p_leaf(
    const BatchSys *__restrict__ sys, const int64_t *__restrict__ loff,
    const int64_t *__restrict__ lvl, int S, const BatchDom *__restrict__ dom,
    const float4 *__restrict__ src, const int32_t *__restrict__ start,
    const int32_t *__restrict__ sp, const int32_t *__restrict__ perm, int ws,
    const float2 *__restrict__ Bc, int p1, double *__restrict__ vx, double *__restrict__ vy,
    int32_t *__restrict__ flag, const int2 *__restrict__ win, int64_t gl, float4 *st, float4 *acc)
{
    const int32_t t0 = start[gl], t1 = start[gl + 1];
    if (t1 <= t0) return;
    const int s = find_seg(loff, S, gl);
    const BatchSys sy = sys[s];
    const uint32_t b = (uint32_t)(gl - sy.leafbase);
    const double sd = dom[s].s;
    const float s_f = (float)sd;
    const float inv_s = 1.0f / s_f;
    int R = near_radius(ws, sd, sy.c2);
    if (R > kRMax) {
        if (threadIdx.x == 0) atomicOr(flag, 1);
        R = kRMax;
    }
    PairParams pp{};
    pp.nlog2e_ks_f = sy.nlog2e_ks_f;
    pp.c2_f = sy.c2_f;
    pp.scale_f = sy.scale_f;
    const BatchNear nf{win, R, ws, (int)compact_bits(b), (int)compact_bits(b >> 1), 1 << sy.L,
                       sd, sy.c2, 1.4426950408889634074 * sy.inv_ks, sy.leafbase, s_f};
    const int nnear = (2 * kRMax + 1) * (2 * kRMax + 1);
    const float2 *Bb = (sy.L >= 2 && Bc) ? Bc + (lvl[sy.L * (S + 1) + s] + b) * p1 : nullptr;
    const int32_t ts = sp[gl];
    const float lam = __builtin_sqrtf(-pp.nlog2e_ks_f);
    auto tpos = [&](int32_t i) {
        const float4 q = src[ts + (i - t0)];
        return make_float2(q.x, q.y);
    };
    auto out = [&](int32_t i, float ha, float la, float hb, float lb) {
        double fr = 0.0, fi = 0.0;
        const float2 tp = tpos(i);
        if (Bb) l2p_f32(Bb, p1, tp.x * inv_s, tp.y * inv_s, fr, fi);
        const int32_t o = perm[i];
        const double qr = -(((double)ha + (double)la) * (double)lam + fr);
        const double qi = ((double)hb + (double)lb) * (double)lam - fi;
        const double two_pi = 2.0 * 3.141592653589793;
        vx[o] = -qi / two_pi;  // _induced.py:334-336
        vy[o] = -qr / two_pi;
    };
    p2p_leaf<KIND, false>(src, sp, st, acc, nnear, nf, t0, t1, tpos, pp, lam, out);
}

// One workgroup of 4 waves per leaf; workgroups are dealt round-robin over
// the 8 XCDs and those of XCD k walk the Morton leaf range [split[k],
// split[k+1]) holding an eighth of the batch's blobs (equal work per XCD,
// empty leaves skipped).
template <int KIND>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void batch_p2p_f32s(
    const BatchSys *__restrict__ sys, const int64_t *__restrict__ loff,
    const int64_t *__restrict__ lvl, int S, const BatchDom *__restrict__ dom,
    const float4 *__restrict__ src, const int32_t *__restrict__ start,
    const int32_t *__restrict__ sp, const int32_t *__restrict__ perm, int ws,
    const float2 *__restrict__ Bc, int p1, double *__restrict__ vx, double *__restrict__ vy,
    int32_t *__restrict__ flag, const int32_t *__restrict__ split, const int2 *__restrict__ win)
{
    __shared__ float4 st[P2P_STAGE];
    __shared__ float4 acc[P2P_CHUNKS * 4 * 16];
    const int32_t hi = split[(blockIdx.x & 7u) + 1];
    for (int32_t gl = split[blockIdx.x & 7u] + (int32_t)(blockIdx.x >> 3); gl < hi;
         gl += (int32_t)(gridDim.x >> 3))
        batch_p2p_leaf<KIND>(sys, loff, lvl, S, dom, src, start, sp, perm, ws, Bc, p1, vx, vy,
                             flag, win, gl, st, acc);
}

template <int KIND>
__global__ __launch_bounds__(64) void batch_p2p_f64(
    const BatchSys *__restrict__ sys, const int64_t *__restrict__ loff,
    const int64_t *__restrict__ lvl, int S, const BatchDom *__restrict__ dom,
    const double4 *__restrict__ src, const int32_t *__restrict__ start,
    const int32_t *__restrict__ perm, int ws, const double2 *__restrict__ Bc, int p1,
    double *__restrict__ vx, double *__restrict__ vy, int32_t *__restrict__ flag)
{
    const int lane = threadIdx.x;
    const int64_t g = blockIdx.x;
    const int32_t t0 = start[g], t1 = start[g + 1];
    if (t1 <= t0) return;
    const int s = find_seg(loff, S, g);
    const BatchSys sy = sys[s];
    const uint32_t b = (uint32_t)(g - sy.leafbase);
    const int bx = (int)compact_bits(b), by = (int)compact_bits(b >> 1);
    const int n = 1 << sy.L;
    const BatchDom d = dom[s];
    int R = near_radius(ws, d.s, sy.c2);
    if (R > kRMax) {
        if (lane == 0) atomicOr(flag, 1);
        R = kRMax;
    }
    PairParams pp{};
    pp.inv_ks = sy.inv_ks;
    pp.c2 = sy.c2;
    pp.scale = sy.scale;
    double cx, cy;
    leaf_center(b, d.x0, d.y0, d.s, cx, cy);
    const double inv_s = 1.0 / d.s;
    const double2 *Bb = (sy.L >= 2 && Bc) ? Bc + (lvl[sy.L * (S + 1) + s] + b) * p1 : nullptr;
    const double two_pi = 2.0 * 3.141592653589793;
    for (int32_t tc = t0; tc < t1; tc += 64) {
        const int32_t i = tc + lane;
        const bool valid = i < t1;
        const double4 tp = valid ? src[i] : make_double4(cx, cy, 0.0, 0.0);
        double a = 0.0, bb = 0.0;
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
                const int sx = bx + dx, syy = by + dy;
                if (sx < 0 || syy < 0 || sx >= n || syy >= n) continue;
                if (!near_pred(dx, dy, ws, d.s, sy.c2)) continue;
                const int64_t sg = sy.leafbase + morton2((uint32_t)sx, (uint32_t)syy);
                for (int32_t j = start[sg]; j < start[sg + 1]; ++j) {
                    const double4 sp = src[j];
                    const double ddx = tp.x - sp.x, ddy = tp.y - sp.y;
                    const double P = pair_f64<KIND>(ddx, ddy, sp.z, pp);
                    a = __builtin_fma(P, ddx, a);
                    bb = __builtin_fma(P, ddy, bb);
                }
            }
        double fr = 0.0, fi = 0.0;
        if (Bb) {
            const double zr = (tp.x - cx) * inv_s, zi = (tp.y - cy) * inv_s;
            for (int k = p1 - 1; k >= 0; --k) {
                const double2 c = Bb[k];
                const double nr = fr * zr - fi * zi + c.x;
                const double ni = fr * zi + fi * zr + c.y;
                fr = nr;
                fi = ni;
            }
        }
        if (valid) {
            const int32_t o = perm[i];
            const double qr = -(a + fr), qi = bb - fi;
            vx[o] = -qi / two_pi;
            vy[o] = -qr / two_pi;
        }
    }
}

// ------------------------------------------------------------- host ------
namespace {

double binom_h(int n, int k)
{
    if (k < 0 || k > n) return 0.0;
    long double r = 1.0L;
    for (int i = 1; i <= k; ++i) r = r * (long double)(n - k + i) / (long double)i;
    return (double)r;
}

template <class T, int P1>
int launch_bm2l(Lane *L, const BatchPlan &P, int l, int blocks, const void *M, void *B,
                const void *U, const void *bin, int32_t *flag, const BatchSys *sys,
                const int64_t *lvl, const BatchDom *dom, const int32_t *start)
{
    using V = typename C2<T>::type;
    batch_m2l<T, P1><<<blocks, 64, 0, L->stream>>>(sys, lvl, P.S, l, dom, start, P.p + 1, P.ws,
                                                   (const V *)M, (V *)B, (const V *)U,
                                                   (const T *)bin, flag);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

template <class T>
int dispatch_bm2l(Lane *L, const BatchPlan &P, int l, int64_t nb, const void *M, void *B,
                  const void *U, const void *bin, int32_t *flag, const BatchSys *sys,
                  const int64_t *lvl, const BatchDom *dom, const int32_t *start)
{
    const int p1 = P.p + 1;
    if (sizeof(T) == 4 && p1 <= BM_ROWS) {
        batch_m2l_mfma<<<(unsigned)nb, 64, 0, L->stream>>>(
            sys, lvl, P.S, l, dom, start, p1, P.ws, (const float2 *)M, (float2 *)B,
            (const float2 *)U, (const float *)bin, flag);
        PHF_HIP(hipGetLastError());
        return PHF_OK;
    }
    const int blocks = (int)((nb + 64 / p1 - 1) / (64 / p1));
#define PHF_BM2L(Q) \
    if (p1 <= Q) return launch_bm2l<T, Q>(L, P, l, blocks, M, B, U, bin, flag, sys, lvl, dom, start);
    PHF_BM2L(9)
    PHF_BM2L(13)
    PHF_BM2L(17)
    PHF_BM2L(21)
    PHF_BM2L(25)
    PHF_BM2L(32)
    PHF_BM2L(41)
#undef PHF_BM2L
    set_error("expansion order %d too large (max 40)", P.p);
    return PHF_ERR_UNSUPPORTED;
}

}  // namespace

int batch_plan(const int64_t *sizes, const double *sigma, int S, int p, int leaf, int ws, int prec,
               double k, BatchPlan *P)
{
    PHF_REQUIRE(S >= 1 && S <= 65536, "system count out of range (%d)", S);
    PHF_REQUIRE(p >= 1 && p <= 40, "order out of range (%d)", p);
    PHF_REQUIRE(leaf >= 1, "leaf must be >= 1");
    P->S = S;
    P->p = p;
    P->ws = ws;
    P->prec = prec;
    P->sys.assign(S, BatchSys{});
    P->poff.assign(S + 1, 0);
    P->loff.assign(S + 1, 0);
    int Lmax = 0;
    for (int s = 0; s < S; ++s) {
        PHF_REQUIRE(sizes[s] >= 0, "negative system size");
        PHF_REQUIRE(std::isfinite(sigma[s]) && sigma[s] > 0, "sigma[%d] must be > 0", s);
        const int L = fmm_level(sizes[s], leaf, 35);
        BatchSys &b = P->sys[s];
        b.off = P->poff[s];
        b.n = (int32_t)sizes[s];
        b.L = L;
        b.leafbase = P->loff[s];
        const double ks = k * sigma[s] * sigma[s];
        const double xopt = std::sqrt(1.2564312086261696770 * ks);
        const PairParams pp = make_oseen_params(xopt, 5.0 * xopt);
        b.cutoff = 5.0 * xopt;
        b.c2 = pp.c2;
        b.inv_ks = pp.inv_ks;
        b.scale = pp.scale;
        b.c2_f = pp.c2_f;
        b.nlog2e_ks_f = pp.nlog2e_ks_f;
        b.scale_f = pp.scale_f;
        b.nocut = pp.f32_nocut;
        P->poff[s + 1] = P->poff[s] + sizes[s];
        P->loff[s + 1] = P->loff[s] + (int64_t(1) << (2 * L));
        Lmax = std::max(Lmax, L);
    }
    P->N = P->poff[S];
    P->Lmax = Lmax;
    P->chunks.clear();
    P->cfirst.assign(S + 1, 0);
    for (int s = 0; s < S; ++s) {
        P->cfirst[s] = (int32_t)P->chunks.size();
        for (int64_t b = P->poff[s]; b < P->poff[s + 1]; b += kBatchChunk)
            P->chunks.push_back(BatchChunk{b, std::min(b + kBatchChunk, P->poff[s + 1]), s, 0});
    }
    P->cfirst[S] = (int32_t)P->chunks.size();
    P->nleaf = P->loff[S];
    PHF_REQUIRE(P->nleaf < (int64_t(1) << 31) && P->N < (int64_t(1) << 31),
                "batch too large (%lld points, %lld leaves)", (long long)P->N,
                (long long)P->nleaf);
    // level-major box bases
    P->lvl.assign((size_t)(Lmax + 1) * (S + 1), 0);
    int64_t base = 0;
    for (int l = 0; l <= Lmax; ++l) {
        for (int s = 0; s < S; ++s) {
            P->lvl[(size_t)l * (S + 1) + s] = base;
            if (P->sys[s].L >= l) base += int64_t(1) << (2 * l);
        }
        P->lvl[(size_t)l * (S + 1) + S] = base;
    }
    P->nbox = base;
    P->nocut_all = true;
    for (auto &b : P->sys) P->nocut_all = P->nocut_all && b.nocut;
    return PHF_OK;
}

int batch_alloc(Lane *L, const BatchPlan &P, BatchBuffers *Bf)
{
    const int p1 = P.p + 1;
    const bool f32 = P.prec == PHF_PREC_F32C;
    const size_t N = (size_t)std::max<int64_t>(P.N, 1);
    PHF_ALLOC(Bf->sys, L->arena, BatchSys, (size_t)P.S);
    PHF_ALLOC(Bf->dom, L->arena, BatchDom, (size_t)P.S);
    PHF_ALLOC(Bf->split, L->arena, int32_t, 9);
    PHF_ALLOC(Bf->chunks, L->arena, BatchChunk, std::max<size_t>(P.chunks.size(), 1));
    PHF_ALLOC(Bf->cfirst, L->arena, int32_t, (size_t)P.S + 1);
    PHF_ALLOC(Bf->cpart, L->arena, char, 32 * std::max<size_t>(P.chunks.size(), 1));
    PHF_ALLOC(Bf->poff, L->arena, int64_t, (size_t)P.S + 1);
    PHF_ALLOC(Bf->loff, L->arena, int64_t, (size_t)P.S + 1);
    PHF_ALLOC(Bf->lvl, L->arena, int64_t, P.lvl.size());
    PHF_ALLOC(Bf->k0, L->arena, uint32_t, N);
    PHF_ALLOC(Bf->k1, L->arena, uint32_t, N);
    PHF_ALLOC(Bf->v0, L->arena, int32_t, N);
    PHF_ALLOC(Bf->v1, L->arena, int32_t, N);
    PHF_ALLOC(Bf->pk, L->arena, double4, N);
    PHF_ALLOC(Bf->start, L->arena, int32_t, (size_t)P.nleaf + 1);
    if (f32) {
        const size_t nl = (size_t)P.nleaf + 1;
        PHF_ALLOC(Bf->src, L->arena, char, (N + 3 * nl) * sizeof(float4));
        PHF_ALLOC(Bf->pcnt, L->arena, uint32_t, nl);
        PHF_ALLOC(Bf->psp, L->arena, uint32_t, nl);
        PHF_ALLOC(Bf->psums, L->arena, uint32_t, (size_t)scan_sums_size((int64_t)nl));
    } else {
        PHF_ALLOC(Bf->src, L->arena, char, N * sizeof(double4));
    }
    const size_t csz = f32 ? sizeof(float2) : sizeof(double2);
    PHF_ALLOC(Bf->M, L->arena, char, (size_t)std::max<int64_t>(P.nbox, 1) * p1 * csz);
    PHF_ALLOC(Bf->B, L->arena, char, (size_t)std::max<int64_t>(P.nbox, 1) * p1 * csz);
    PHF_ALLOC(Bf->flag, L->arena, int32_t, 1);
    // constant tables: M2M/L2L shifts, M2L binomials and powers u^k over
    // the interaction-offset window
    const bool tc64 = !f32 || P.p > 16;
    const size_t tsz = tc64 ? sizeof(double2) : sizeof(float2);
    std::vector<char> h_m2m((size_t)4 * p1 * p1 * tsz), h_l2l((size_t)4 * p1 * p1 * tsz);
    for (int c = 0; c < 4; ++c) {
        std::complex<double> d((c & 1) * 0.5 - 0.25, (c >> 1) * 0.5 - 0.25);
        for (int m = 0; m < p1; ++m)
            for (int q = 0; q < p1; ++q) {
                std::complex<double> v1 = 0, v2 = 0;
                if (q <= m) v1 = binom_h(m, q) * std::ldexp(1.0, -q) * std::pow(d, m - q);
                if (q >= m) v2 = binom_h(q, m) * std::pow(d, q - m) * std::ldexp(1.0, -m);
                const size_t at = ((size_t)c * p1 + m) * p1 + q;
                if (tc64) {
                    reinterpret_cast<double2 *>(h_m2m.data())[at] = double2{v1.real(), v1.imag()};
                    reinterpret_cast<double2 *>(h_l2l.data())[at] = double2{v2.real(), v2.imag()};
                } else {
                    reinterpret_cast<float2 *>(h_m2m.data())[at] =
                        float2{(float)v1.real(), (float)v1.imag()};
                    reinterpret_cast<float2 *>(h_l2l.data())[at] =
                        float2{(float)v2.real(), (float)v2.imag()};
                }
            }
    }
    std::vector<char> h_bin((size_t)p1 * p1 * (f32 ? 4 : 8));
    for (int m = 0; m < p1; ++m)
        for (int q = 0; q < p1; ++q) {
            const double v = binom_h(q + m, q);
            if (f32)
                reinterpret_cast<float *>(h_bin.data())[m * p1 + q] = (float)v;
            else
                reinterpret_cast<double *>(h_bin.data())[m * p1 + q] = v;
        }
    const int W2 = 2 * kRU + 1;
    std::vector<char> h_U((size_t)W2 * W2 * (p1 + 1) * csz);
    for (int oy = -kRU; oy <= kRU; ++oy)
        for (int ox = -kRU; ox <= kRU; ++ox) {
            std::complex<double> u = 0.0;
            if (ox || oy) u = -1.0 / std::complex<double>(ox, oy);
            std::complex<double> pw = 1.0;
            for (int q = 0; q <= p1; ++q) {
                const size_t at = ((size_t)(oy + kRU) * W2 + (ox + kRU)) * (p1 + 1) + q;
                if (f32)
                    reinterpret_cast<float2 *>(h_U.data())[at] =
                        float2{(float)pw.real(), (float)pw.imag()};
                else
                    reinterpret_cast<double2 *>(h_U.data())[at] = double2{pw.real(), pw.imag()};
                pw *= u;
            }
        }
    PHF_ALLOC(Bf->m2m, L->arena, char, h_m2m.size());
    PHF_ALLOC(Bf->l2l, L->arena, char, h_l2l.size());
    PHF_ALLOC(Bf->bin, L->arena, char, h_bin.size());
    PHF_ALLOC(Bf->U, L->arena, char, h_U.size());
    // synchronous uploads: the plan is built once per batch
    PHF_HIP(hipMemcpy(Bf->sys, P.sys.data(), sizeof(BatchSys) * P.S, hipMemcpyHostToDevice));
    PHF_HIP(hipMemcpy(Bf->poff, P.poff.data(), sizeof(int64_t) * (P.S + 1), hipMemcpyHostToDevice));
    PHF_HIP(hipMemcpy(Bf->loff, P.loff.data(), sizeof(int64_t) * (P.S + 1), hipMemcpyHostToDevice));
    PHF_HIP(hipMemcpy(Bf->lvl, P.lvl.data(), sizeof(int64_t) * P.lvl.size(), hipMemcpyHostToDevice));
    if (!P.chunks.empty())
        PHF_HIP(hipMemcpy(Bf->chunks, P.chunks.data(), sizeof(BatchChunk) * P.chunks.size(),
                          hipMemcpyHostToDevice));
    PHF_HIP(hipMemcpy(Bf->cfirst, P.cfirst.data(), sizeof(int32_t) * (P.S + 1),
                      hipMemcpyHostToDevice));
    PHF_HIP(hipMemcpy(Bf->m2m, h_m2m.data(), h_m2m.size(), hipMemcpyHostToDevice));
    PHF_HIP(hipMemcpy(Bf->l2l, h_l2l.data(), h_l2l.size(), hipMemcpyHostToDevice));
    PHF_HIP(hipMemcpy(Bf->bin, h_bin.data(), h_bin.size(), hipMemcpyHostToDevice));
    PHF_HIP(hipMemcpy(Bf->U, h_U.data(), h_U.size(), hipMemcpyHostToDevice));
    // near-window offsets |d| <= kRMax sorted by box gap (then raster order)
    {
        std::vector<int2> win;
        for (int dy = -kRMax; dy <= kRMax; ++dy)
            for (int dx = -kRMax; dx <= kRMax; ++dx) win.push_back(make_int2(dx, dy));
        auto gap2 = [](const int2 &o) {
            const int ax = std::max(std::abs(o.x) - 1, 0), ay = std::max(std::abs(o.y) - 1, 0);
            return ax * ax + ay * ay;
        };
        std::stable_sort(win.begin(), win.end(),
                         [&](const int2 &a, const int2 &b) { return gap2(a) < gap2(b); });
        PHF_ALLOC(Bf->win, L->arena, int2, win.size());
        PHF_HIP(hipMemcpy(Bf->win, win.data(), sizeof(int2) * win.size(), hipMemcpyHostToDevice));
    }
    // radix-sort scratch (fixed for the plan)
    const int64_t rw = radix_ws_size(P.N);
    PHF_ALLOC(Bf->hist, L->arena, uint32_t, (size_t)rw);
    PHF_ALLOC(Bf->offs, L->arena, uint32_t, (size_t)rw);
    return PHF_OK;
}

template <class T, class TC>
static int batch_enqueue_typed(Lane *L, const BatchPlan &P, BatchBuffers &Bf, const double *x,
                               const double *y, const double *w, double *vx, double *vy)
{
    using V = typename C2<T>::type;
    using VC = typename C2<TC>::type;
    const int S = P.S, p1 = P.p + 1;
    const int64_t N = P.N;
    batch_clear_flag<<<1, 1, 0, L->stream>>>(Bf.flag);
    if (!P.chunks.empty())
        batch_bbox_part<<<(unsigned)P.chunks.size(), 256, 0, L->stream>>>(
            Bf.chunks, x, y, reinterpret_cast<double4 *>(Bf.cpart));
    batch_domain<<<S, 64, 0, L->stream>>>(Bf.sys, Bf.cfirst,
                                          reinterpret_cast<const double4 *>(Bf.cpart), P.ws,
                                          P.p + 1, P.prec == PHF_PREC_F32C ? 16 : 64, Bf.dom);
    if (N > 0)
        batch_keys<<<div_up(N, 256), 256, 0, L->stream>>>(Bf.sys, Bf.poff, S, Bf.dom, x, y, w, N,
                                                          Bf.k0, Bf.v0, Bf.pk);
    PHF_HIP(hipGetLastError());
    int key_bits = 1;
    while (key_bits < 32 && (int64_t(1) << key_bits) < P.nleaf) ++key_bits;
    uint32_t *ks;
    int32_t *vs;
    PHF_TRY(device_radix_sort_ws(L, Bf.k0, Bf.v0, Bf.k1, Bf.v1, N, key_bits, Bf.hist, Bf.offs,
                                 &ks, &vs));
    PHF_TRY(leaf_starts_device(L, ks, N, P.nleaf, Bf.start));
    const int32_t *sp = Bf.start;  // source ranges per leaf (padded for float32)
    if (sizeof(T) == 4) {
        PHF_TRY(pad_source_ranges(L, Bf.start, P.nleaf, Bf.pcnt, Bf.psp, Bf.psums));
        sp = reinterpret_cast<const int32_t *>(Bf.psp);
        PHF_TRY(zero_source_pads(L, Bf.start, sp, P.nleaf, reinterpret_cast<float4 *>(Bf.src)));
    }
    if (N > 0)
        batch_gather<T><<<div_up(N, 256), 256, 0, L->stream>>>(Bf.sys, Bf.loff, S, Bf.dom, Bf.pk,
                                                               ks, vs, Bf.start, sp, N, Bf.src);
    PHF_HIP(hipGetLastError());
    V *M = reinterpret_cast<V *>(Bf.M), *B = reinterpret_cast<V *>(Bf.B);
    const VC *m2m = reinterpret_cast<const VC *>(Bf.m2m);
    const VC *l2l = reinterpret_cast<const VC *>(Bf.l2l);
    if (P.Lmax >= 2) {
        batch_p2m<T><<<(unsigned)P.nleaf, 64, (size_t)64 * p1 * sizeof(V), L->stream>>>(
            Bf.sys, Bf.loff, Bf.lvl, S, Bf.dom, Bf.src, sp, p1, M);
        for (int l = P.Lmax - 1; l >= 2; --l) {
            const int64_t nb = P.lvl[(size_t)l * (S + 1) + S] - P.lvl[(size_t)l * (S + 1)];
            if (nb > 0)
                batch_m2m<T, TC><<<div_up(nb * p1, 256), 256, 0, L->stream>>>(
                    Bf.sys, Bf.lvl, S, l, Bf.start, p1, M, m2m);
        }
        PHF_HIP(hipGetLastError());
        for (int l = 2; l <= P.Lmax; ++l) {
            const int64_t nb = P.lvl[(size_t)l * (S + 1) + S] - P.lvl[(size_t)l * (S + 1)];
            if (nb > 0)
                PHF_TRY(dispatch_bm2l<T>(L, P, l, nb, M, B, Bf.U, Bf.bin,
                                         Bf.flag, Bf.sys, Bf.lvl, Bf.dom, Bf.start));
        }
        for (int l = 2; l < P.Lmax; ++l) {
            const int64_t nb =
                P.lvl[(size_t)(l + 1) * (S + 1) + S] - P.lvl[(size_t)(l + 1) * (S + 1)];
            if (nb > 0)
                batch_l2l<T, TC><<<div_up(nb * p1, 256), 256, 0, L->stream>>>(
                    Bf.sys, Bf.lvl, S, l, Bf.start, p1, B, l2l);
        }
        PHF_HIP(hipGetLastError());
    }
    const V *Bc = P.Lmax >= 2 ? B : nullptr;
    if (P.nleaf > 0) {
        if (sizeof(T) == 4) {
            const uint32_t nl8 = (uint32_t)((P.nleaf + 7) / 8);
            PHF_TRY(xcd_split(L, Bf.start, P.nleaf, Bf.split));
            if (P.nocut_all)
                batch_p2p_f32s<PAIR_OSEEN_NOCUT><<<8 * nl8, 256, 0, L->stream>>>(
                    Bf.sys, Bf.loff, Bf.lvl, S, Bf.dom, (const float4 *)Bf.src, Bf.start, sp, vs,
                    P.ws, (const float2 *)Bc, p1, vx, vy, Bf.flag, Bf.split, Bf.win);
            else
                batch_p2p_f32s<PAIR_OSEEN><<<8 * nl8, 256, 0, L->stream>>>(
                    Bf.sys, Bf.loff, Bf.lvl, S, Bf.dom, (const float4 *)Bf.src, Bf.start, sp, vs,
                    P.ws, (const float2 *)Bc, p1, vx, vy, Bf.flag, Bf.split, Bf.win);
        } else {
            batch_p2p_f64<PAIR_OSEEN><<<(unsigned)P.nleaf, 64, 0, L->stream>>>(
                Bf.sys, Bf.loff, Bf.lvl, S, Bf.dom, (const double4 *)Bf.src, Bf.start, vs, P.ws,
                (const double2 *)Bc, p1, vx, vy, Bf.flag);
        }
    }
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

int batch_enqueue(Lane *L, const BatchPlan &P, BatchBuffers &Bf, const double *x, const double *y,
                  const double *w, double *vx, double *vy)
{
    if (P.prec == PHF_PREC_F32C) {
        if (P.p > 16) return batch_enqueue_typed<float, double>(L, P, Bf, x, y, w, vx, vy);
        return batch_enqueue_typed<float, float>(L, P, Bf, x, y, w, vx, vy);
    }
    return batch_enqueue_typed<double, double>(L, P, Bf, x, y, w, vx, vy);
}

}  // namespace phf

// ------------------------------------------------------------ C entry ----
using namespace phf;

struct phf_batch_s {
    std::mutex mu;
    Lane lane;
    BatchPlan plan;
    BatchBuffers buf;
    double *x = nullptr, *y = nullptr, *w = nullptr, *vx = nullptr, *vy = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<double> sigma;
    // convection + remesh (phf_batch_step)
    bool have_remesh = false;
    BatchRemeshPlan rplan;
    BatchRemeshBuffers rbuf;
    hipGraph_t graph_s = nullptr;
    hipGraphExec_t exec_s = nullptr;
    std::vector<int64_t> mo;  // (offset, count) per system after the last step
    std::map<int, std::vector<double>> spill;  // x|y|w of systems remeshed apart
};

static void drop_step_graph(phf_batch *b)
{
    if (b->exec_s) (void)hipGraphExecDestroy(b->exec_s);
    if (b->graph_s) (void)hipGraphDestroy(b->graph_s);
    b->exec_s = nullptr;
    b->graph_s = nullptr;
}

extern "C" {

int phf_batch_destroy(phf_batch *b)
{
    clear_error();
    if (!b) return PHF_OK;
    {
        std::lock_guard<std::mutex> lk(b->mu);
        if (b->lane.stream) (void)hipStreamSynchronize(b->lane.stream);
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        drop_step_graph(b);
        fmm_graphs_release(&b->lane);
        b->rbuf.arena.release();
        b->lane.arena.release();
        b->lane.pin.release();
        for (auto &ev : b->lane.ev)
            if (ev) (void)hipEventDestroy(ev);
        if (b->lane.stream) (void)hipStreamDestroy(b->lane.stream);
    }
    delete b;
    return PHF_OK;
}

int phf_batch_create(const int64_t *sizes, const double *sigma, int32_t count,
                     const phf_batch_params *bp, phf_batch **out)
{
    clear_error();
    PHF_REQUIRE(out && sizes && sigma && bp, "NULL argument");
    *out = nullptr;
    PHF_REQUIRE(bp->prec == PHF_PREC_F64 || bp->prec == PHF_PREC_F32C, "unknown precision %d",
                bp->prec);
    PHF_REQUIRE(std::isfinite(bp->k) && bp->k > 0, "k must be > 0");
    PHF_TRY(Context::get().init());
    PHF_HIP(hipSetDevice(Context::get().device()));
    phf_batch *b = new phf_batch();
    b->sigma.assign(sigma, sigma + std::max(count, 0));
    int rc = [&]() -> int {
        PHF_TRY(batch_plan(sizes, sigma, count, bp->order > 0 ? bp->order : 16,
                           bp->leaf > 0 ? bp->leaf : 64, bp->ws > 0 ? bp->ws : 1, bp->prec, bp->k,
                           &b->plan));
        PHF_HIP(hipStreamCreateWithFlags(&b->lane.stream, hipStreamNonBlocking));
        for (auto &ev : b->lane.ev) PHF_HIP(hipEventCreate(&ev));
        Lane *L = &b->lane;
        const size_t N = (size_t)std::max<int64_t>(b->plan.N, 1);
        PHF_ALLOC(b->x, L->arena, double, N);
        PHF_ALLOC(b->y, L->arena, double, N);
        PHF_ALLOC(b->w, L->arena, double, N);
        PHF_ALLOC(b->vx, L->arena, double, N);
        PHF_ALLOC(b->vy, L->arena, double, N);
        PHF_TRY(batch_alloc(L, b->plan, &b->buf));
        return PHF_OK;
    }();
    if (rc != PHF_OK) {
        phf_batch_destroy(b);
        return rc;
    }
    *out = b;
    return PHF_OK;
}

int phf_batch_set_system(phf_batch *b, int32_t s, const double *x, const double *y,
                         const double *w)
{
    clear_error();
    PHF_REQUIRE(b != nullptr, "batch is NULL");
    PHF_REQUIRE(s >= 0 && s < b->plan.S, "system index %d out of range", s);
    const int64_t n = b->plan.sys[s].n, o = b->plan.sys[s].off;
    PHF_TRY(check_host_array(x, n, "x"));
    PHF_TRY(check_host_array(y, n, "y"));
    PHF_TRY(check_host_array(w, n, "w"));
    std::lock_guard<std::mutex> lk(b->mu);
    PHF_HIP(hipSetDevice(Context::get().device()));
    if (n > 0) {
        PHF_HIP(hipMemcpyAsync(b->x + o, x, sizeof(double) * n, hipMemcpyHostToDevice, b->lane.stream));
        PHF_HIP(hipMemcpyAsync(b->y + o, y, sizeof(double) * n, hipMemcpyHostToDevice, b->lane.stream));
        PHF_HIP(hipMemcpyAsync(b->w + o, w, sizeof(double) * n, hipMemcpyHostToDevice, b->lane.stream));
    }
    PHF_HIP(hipStreamSynchronize(b->lane.stream));
    return PHF_OK;
}

int phf_batch_velocity(phf_batch *b, int32_t use_graph, phf_batch_info *info)
{
    clear_error();
    PHF_REQUIRE(b != nullptr, "batch is NULL");
    std::lock_guard<std::mutex> lk(b->mu);
    PHF_HIP(hipSetDevice(Context::get().device()));
    Lane *L = &b->lane;
    hipEvent_t e0 = L->ev[0], e1 = L->ev[1];
    if (use_graph && !b->exec) {
        PHF_HIP(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_enqueue(L, b->plan, b->buf, b->x, b->y, b->w, b->vx, b->vy);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(L->stream, &g);
        if (rc != PHF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        PHF_HIP(e);
        b->graph = g;
        PHF_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    }
    PHF_HIP(hipEventRecord(e0, L->stream));
    if (use_graph)
        PHF_HIP(hipGraphLaunch(b->exec, L->stream));
    else
        PHF_TRY(batch_enqueue(L, b->plan, b->buf, b->x, b->y, b->w, b->vx, b->vy));
    PHF_HIP(hipEventRecord(e1, L->stream));
    int32_t flag = 0;
    PHF_HIP(hipMemcpyAsync(&flag, b->buf.flag, sizeof(int32_t), hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (info) {
        info->ms = ms;
        info->graph = use_graph ? 1 : 0;
        info->flags = flag;
        info->total_points = b->plan.N;
        info->levels_max = b->plan.Lmax;
    }
    if (flag) {
        set_error("a system's smoothing cutoff spans more than %d leaf boxes; use a larger leaf "
                  "size", 6);
        return PHF_ERR_UNSUPPORTED;
    }
    return PHF_OK;
}

int phf_batch_get_velocity(phf_batch *b, int32_t s, double *vx, double *vy)
{
    clear_error();
    PHF_REQUIRE(b != nullptr, "batch is NULL");
    PHF_REQUIRE(s >= 0 && s < b->plan.S, "system index %d out of range", s);
    std::lock_guard<std::mutex> lk(b->mu);
    const int64_t n = b->plan.sys[s].n, o = b->plan.sys[s].off;
    PHF_REQUIRE(n == 0 || (vx && vy), "output arrays are NULL");
    if (n > 0) {
        PHF_HIP(hipMemcpyAsync(vx, b->vx + o, sizeof(double) * n, hipMemcpyDeviceToHost, b->lane.stream));
        PHF_HIP(hipMemcpyAsync(vy, b->vy + o, sizeof(double) * n, hipMemcpyDeviceToHost, b->lane.stream));
    }
    PHF_HIP(hipStreamSynchronize(b->lane.stream));
    return PHF_OK;
}

int phf_batch_set_remesh(phf_batch *b, const phf_batch_remesh_params *rp)
{
    clear_error();
    PHF_REQUIRE(b && rp, "NULL argument");
    PHF_REQUIRE(std::isfinite(rp->dt) && std::isfinite(rp->vinf_x) && std::isfinite(rp->vinf_y),
                "dt and vinf must be finite");
    PHF_REQUIRE(std::isfinite(rp->overlap) && rp->overlap > 0, "overlap must be > 0");
    PHF_REQUIRE(std::isfinite(rp->x0h) && std::isfinite(rp->y0h) && std::isfinite(rp->c),
                "lattice origin and c must be finite");
    PHF_REQUIRE(std::isfinite(rp->cell_factor) && rp->cell_factor <= 64.0,
                "cell_factor must be <= 64");
    std::lock_guard<std::mutex> lk(b->mu);
    PHF_HIP(hipSetDevice(Context::get().device()));
    PHF_HIP(hipStreamSynchronize(b->lane.stream));
    BatchRemeshPlan R;
    PHF_TRY(batch_remesh_plan(b->plan, b->sigma.data(), *rp, &R));
    drop_step_graph(b);
    b->have_remesh = false;
    b->rbuf.arena.release();
    b->rbuf = BatchRemeshBuffers();
    PHF_TRY(batch_remesh_alloc(b->plan, R, &b->rbuf));
    b->rplan = std::move(R);
    b->mo.assign((size_t)2 * b->plan.S, 0);
    b->spill.clear();
    b->have_remesh = true;
    return PHF_OK;
}

int phf_batch_step(phf_batch *b, int32_t use_graph, phf_batch_info *info)
{
    clear_error();
    PHF_REQUIRE(b != nullptr, "batch is NULL");
    std::lock_guard<std::mutex> lk(b->mu);
    PHF_REQUIRE(b->have_remesh, "phf_batch_set_remesh has not been called");
    PHF_HIP(hipSetDevice(Context::get().device()));
    Lane *L = &b->lane;
    hipEvent_t e0 = L->ev[0], e1 = L->ev[1];
    auto enqueue = [&]() -> int {
        PHF_TRY(batch_enqueue(L, b->plan, b->buf, b->x, b->y, b->w, b->vx, b->vy));
        return batch_remesh_enqueue(L, b->plan, b->buf, b->rplan, b->rbuf, b->x, b->y, b->w,
                                    b->vx, b->vy);
    };
    if (use_graph && !b->exec_s) {
        PHF_HIP(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(L->stream, &g);
        if (rc != PHF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        PHF_HIP(e);
        b->graph_s = g;
        PHF_HIP(hipGraphInstantiate(&b->exec_s, g, nullptr, nullptr, 0));
    }
    PHF_HIP(hipEventRecord(e0, L->stream));
    if (use_graph)
        PHF_HIP(hipGraphLaunch(b->exec_s, L->stream));
    else
        PHF_TRY(enqueue());
    PHF_HIP(hipEventRecord(e1, L->stream));
    const int S = b->plan.S;
    PHF_TRY(L->pin.ensure(sizeof(int64_t) * (2 * (size_t)S + 1)));
    int64_t *hp = reinterpret_cast<int64_t *>(L->pin.ptr);
    PHF_HIP(hipMemcpyAsync(hp, b->rbuf.mo, sizeof(int64_t) * 2 * S, hipMemcpyDeviceToHost,
                           L->stream));
    PHF_HIP(hipMemcpyAsync(hp + 2 * S, b->buf.flag, sizeof(int32_t), hipMemcpyDeviceToHost,
                           L->stream));
    PHF_TRY(L->pin.mark(L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const int32_t flag = *reinterpret_cast<int32_t *>(hp + 2 * S);
    b->mo.assign(hp, hp + 2 * S);
    if (flag) {
        set_error("a system's smoothing cutoff spans more than %d leaf boxes; use a larger leaf "
                  "size", 6);
        return PHF_ERR_UNSUPPORTED;
    }
    // systems whose lattice range did not fit: single-system remesh
    b->spill.clear();
    int32_t spilled = 0;
    int64_t total = 0;
    for (int s = 0; s < S; ++s) {
        if (b->mo[2 * s + 1] >= 0) {
            total += b->mo[2 * s + 1];
            continue;
        }
        ++spilled;
        LaneGuard g;
        PHF_TRY(g.rc);
        Lane *T = g.lane;
        PHF_TRY(reset_arena(T));
        const BatchRm &r = b->rplan.rm[s];
        const int64_t o = b->plan.sys[s].off, n = b->plan.sys[s].n;
        const double *xs = b->rplan.p.dt != 0.0 ? b->rbuf.xn : b->x;
        const double *ys = b->rplan.p.dt != 0.0 ? b->rbuf.yn : b->y;
        RemeshOut ro;
        PHF_TRY(remesh_device(T, xs + o, ys + o, b->w + o, n, r.h, 1.0, r.x0, r.y0,
                              b->rplan.p.c, &ro));
        std::vector<double> &v = b->spill[s];
        v.resize((size_t)3 * ro.m);
        if (ro.m > 0) {
            PHF_HIP(hipMemcpyAsync(v.data(), ro.x, sizeof(double) * ro.m, hipMemcpyDeviceToHost,
                                   T->stream));
            PHF_HIP(hipMemcpyAsync(v.data() + ro.m, ro.y, sizeof(double) * ro.m,
                                   hipMemcpyDeviceToHost, T->stream));
            PHF_HIP(hipMemcpyAsync(v.data() + 2 * ro.m, ro.w, sizeof(double) * ro.m,
                                   hipMemcpyDeviceToHost, T->stream));
        }
        PHF_HIP(hipStreamSynchronize(T->stream));
        b->mo[2 * s + 1] = ro.m;
        total += ro.m;
    }
    if (info) {
        info->ms = ms;
        info->graph = use_graph ? 1 : 0;
        info->flags = flag;
        info->total_points = b->plan.N;
        info->levels_max = b->plan.Lmax;
        info->remeshed = total;
        info->spilled = spilled;
    }
    return PHF_OK;
}

int phf_batch_get_remeshed(phf_batch *b, int32_t s, int64_t *m, double *x, double *y, double *w)
{
    clear_error();
    PHF_REQUIRE(b && m, "NULL argument");
    PHF_REQUIRE(s >= 0 && s < b->plan.S, "system index %d out of range", s);
    std::lock_guard<std::mutex> lk(b->mu);
    PHF_REQUIRE(b->have_remesh && !b->mo.empty(), "no remeshed state (call phf_batch_step)");
    const int64_t cnt = b->mo[2 * s + 1];
    *m = cnt;
    if (!x && !y && !w) return PHF_OK;
    PHF_REQUIRE(cnt == 0 || (x && y && w), "output arrays are NULL");
    if (cnt == 0) return PHF_OK;
    auto it = b->spill.find(s);
    if (it != b->spill.end()) {
        std::memcpy(x, it->second.data(), sizeof(double) * cnt);
        std::memcpy(y, it->second.data() + cnt, sizeof(double) * cnt);
        std::memcpy(w, it->second.data() + 2 * cnt, sizeof(double) * cnt);
        return PHF_OK;
    }
    const int64_t o = b->mo[2 * s];
    PHF_HIP(hipSetDevice(Context::get().device()));
    PHF_HIP(hipMemcpyAsync(x, b->rbuf.xo + o, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                           b->lane.stream));
    PHF_HIP(hipMemcpyAsync(y, b->rbuf.yo + o, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                           b->lane.stream));
    PHF_HIP(hipMemcpyAsync(w, b->rbuf.wo + o, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                           b->lane.stream));
    PHF_HIP(hipStreamSynchronize(b->lane.stream));
    return PHF_OK;
}

}  // extern "C"
