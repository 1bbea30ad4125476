// phf_fmm_dev.hpp -- device helpers shared by the single-system and the
// batched FMM kernels.
#pragma once

#include "phf_device.hpp"

namespace phf {

template <class T> struct C2;
template <> struct C2<float> { using type = float2; };
template <> struct C2<double> { using type = double2; };

template <class T> __device__ __forceinline__ typename C2<T>::type cmk(T r, T i)
{
    return typename C2<T>::type{r, i};
}
template <class V> __device__ __forceinline__ V cmul(V a, V b)
{
    return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// centre of the leaf box with Morton key `key` (side s, origin x0, y0)
__device__ __forceinline__ void leaf_center(uint32_t key, double x0, double y0, double s,
                                            double &cx, double &cy)
{
    uint32_t ix = compact_bits(key), iy = compact_bits(key >> 1);
    cx = x0 + ((double)ix + 0.5) * s;
    cy = y0 + ((double)iy + 0.5) * s;
}

// box counts at level l from leaf offsets (shift = 2 (L - l))
__device__ __forceinline__ int box_count(const int32_t *__restrict__ start, uint32_t b, int shift)
{
    return start[(int64_t)(b + 1) << shift] - start[(int64_t)b << shift];
}

// ------------------------------------------------------ domain choice --
// Shared by the host (single system) and the device (batched systems) so
// that both build the same tree; contraction is off so that both round
// alike.

// side of a square domain covering the extent ext (at least min_side),
// rounded up to 5 significant bits (D / 2^L is then exact)
__host__ __device__ inline double domain_side(double ext, double min_side)
{
#pragma clang fp contract(off)
    double e = ext > min_side ? ext : min_side;
    if (!(e > 0.0)) e = 1.0;
    int k = 0;
    frexp(e, &k);  // e = f 2^k, f in [0.5, 1)
    const double unit = ldexp(1.0, k - 5);
    return ceil(e / unit) * unit;
}

// leaf offset (dx, dy) is near: within ws boxes, or box gap below the cutoff
__host__ __device__ inline bool near_offset(int dx, int dy, int ws, double s, double c2)
{
#pragma clang fp contract(off)
    const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy;
    if (ax <= ws && ay <= ws) return true;
    const double gx = (double)(ax > 1 ? ax - 1 : 0) * s;
    const double gy = (double)(ay > 1 ? ay - 1 : 0) * s;
    return gx * gx + gy * gy < c2;
}

// near offset whose pairs may lie inside the smoothing core: box gap g with
// shape g^2 log2(e) < 26 (2% margin for float32 rounding of the point
// offsets); beyond it exp(-shape r2) <= 2^-26 for every pair
__host__ __device__ inline bool near_offset_core(int dx, int dy, double s, double shape_log2)
{
#pragma clang fp contract(off)
    const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy;
    const double gx = (double)(ax > 1 ? ax - 1 : 0) * s * 0.99;
    const double gy = (double)(ay > 1 ? ay - 1 : 0) * s * 0.99;
    return shape_log2 * (gx * gx + gy * gy) < 26.0;
}

// Work estimate of the near field and the leaf-level M2L of a tree with
// leaf level L >= 1 and domain side D over extents (ex, ey), nb sources and
// nt targets spread evenly: near pairs (one inside the smoothing core
// weighs 1.2 for its exponential), lanes left idle by a leaf's partial
// target chunk (half a chunk per leaf), and M2L links at p1^2 pairs each
// (4 |near(L-1)| - |near(L)| per occupied leaf).
__host__ __device__ inline double leaf_work(double ex, double ey, double D, int L, double nb,
                                            double nt, double cutoff, int ws, double shape_log2,
                                            int p1, int chunk)
{
#pragma clang fp contract(off)
    const double c2 = cutoff * cutoff;
    const double s = D / ldexp(1.0, L);
    double wnear = 0.0, nl = 0.0, np = 0.0;
    int R = (int)ceil(cutoff / s) + 2;
    R = R > ws ? R : ws;
    for (int dy = -R; dy <= R; ++dy)
        for (int dx = -R; dx <= R; ++dx)
            if (near_offset(dx, dy, ws, s, c2)) {
                nl += 1.0;
                wnear += near_offset_core(dx, dy, s, shape_log2) ? 1.2 : 1.0;
            }
    const double sp = 2.0 * s;
    int Rp = (int)ceil(cutoff / sp) + 2;
    Rp = Rp > ws ? Rp : ws;
    for (int dy = -Rp; dy <= Rp; ++dy)
        for (int dx = -Rp; dx <= Rp; ++dx)
            if (near_offset(dx, dy, ws, sp, c2)) np += 1.0;
    const double n = ldexp(1.0, L);
    double ox = ceil(ex / s), oy = ceil(ey / s);
    ox = ox < 1.0 ? 1.0 : (ox > n ? n : ox);
    oy = oy < 1.0 ? 1.0 : (oy > n ? n : oy);
    const double occ = ox * oy;
    const double spl = nb / occ, tpl = nt / occ;
    const double idle = tpl > 0.0 ? (tpl + 0.5 * chunk) / tpl : 1.0;
    return nt * spl * wnear * idle + occ * (4.0 * np - nl) * (double)(p1 * p1);
}

// Domain side at leaf level L: the bounding-box side, or one enlarged so
// that leaves are at least the Oseen cutoff wide (the near field is then
// the 3 x 3 block of leaves) when leaf_work rates it cheaper.  leaf_side
// overrides the choice: < 0 the bounding box, > 0 leaves at least
// leaf_side * cutoff wide (an autotuned setting); 0 the cost model.
__host__ __device__ inline double choose_domain_side(double ex, double ey, int L, double nb,
                                                     double nt, double cutoff, int ws,
                                                     double shape_log2, int p1, int chunk,
                                                     double leaf_side = 0.0)
{
#pragma clang fp contract(off)
    const double ext = ex > ey ? ex : ey;
    const double D = domain_side(ext, 0.0);
    if (L < 1 || leaf_side < 0.0) return D;
    if (leaf_side > 0.0)
        return domain_side(ext, leaf_side * cutoff * ldexp(1.0, L) * (1.0 + 1e-12));
    const double Dc = domain_side(ext, cutoff * ldexp(1.0, L) * (1.0 + 1e-12));
    if (!(Dc > D)) return D;
    const double w0 = leaf_work(ex, ey, D, L, nb, nt, cutoff, ws, shape_log2, p1, chunk);
    const double w1 = leaf_work(ex, ey, Dc, L, nb, nt, cutoff, ws, shape_log2, p1, chunk);
    return w1 < w0 ? Dc : D;
}

}  // namespace phf
