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

}  // namespace phf
