// phf_m4_dev.hpp -- M4' weight and base-cell index shared by the single-system
// and the batched remesh (same float64 operation order in both, contraction
// off, so both produce bit-identical cells and weights).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace phf {

__device__ __forceinline__ double cube_rn(double a)
{
#pragma clang fp contract(off)
    // a^3 rounded once: a*a as (p + e) exactly, then (p + e) * a.
    const double p = a * a;
    const double e = __builtin_fma(a, a, -p);
    const double t = p * a;
    const double te = __builtin_fma(p, a, -t);
    return t + (te + e * a);
}

// M4' (_interpolation_kernels_1d.cpp:76-87), same operation order.
__host__ __device__ __forceinline__ double m4p_dev(double x, double c, double c2)
{
#pragma clang fp contract(off)
    const double ax = fabs(x);
    if ((ax > 1.0) && (ax < 2.0)) {
        const double t = 2.0 - ax;
        const double t2 = t * t;
        return 0.5 * t2 * (1.0 - ax) - c2 * t2 * (1.0 - 2.0 * ax);
    } else if ((ax >= 0.0) && (ax < 1.0)) {
#ifdef __HIP_DEVICE_COMPILE__
        const double a3 = cube_rn(ax);
#else
        const double a3 = std::pow(ax, 3.0);
#endif
        const double x2 = x * x;
        return 1.0 - 2.5 * x2 + 1.5 * a3 - c2 * (2.0 - 9.0 * x2 + 6.0 * a3);
    }
    return 0.0;
}

// m4p_dev with c = 0 (plain M4', no diffusion term): the c2 products are
// then +-0, and subtracting +-0 leaves the first part unchanged, since that
// part is never -0 (it is < 0 on 1 < |x| < 2, and on |x| < 1 a sum that
// cancels exactly rounds to +0).  Same values as m4p_dev(x, 0, 0), with
// about a third fewer operations.
__host__ __device__ __forceinline__ double m4p_dev0(double x)
{
#pragma clang fp contract(off)
    const double ax = fabs(x);
    if ((ax > 1.0) && (ax < 2.0)) {
        const double t = 2.0 - ax;
        const double t2 = t * t;
        return 0.5 * t2 * (1.0 - ax);
    } else if ((ax >= 0.0) && (ax < 1.0)) {
#ifdef __HIP_DEVICE_COMPILE__
        const double a3 = cube_rn(ax);
#else
        const double a3 = std::pow(ax, 3.0);
#endif
        const double x2 = x * x;
        return 1.0 - 2.5 * x2 + 1.5 * a3;
    }
    return 0.0;
}

// iL = int64(ceil(((x - x0) - 2h + h/2)/h)) - 1  (_regridding.py:157-158)
__device__ __forceinline__ int64_t base_index(double x, double x0, double h)
{
#pragma clang fp contract(off)
    const double h2 = 2.0 * h, hh = 0.5 * h;
    return (int64_t)ceil((((x - x0) - h2) + hh) / h) - 1;
}

// Tutty's diffusion weight of node i - 1 + a (a = 0..3) for a blob at
// delta = (x - x_i)/h from the node right of its base cell, c2 = (hNu/h)^2
// (_evolution.py:385-405, same operation order).
__host__ __device__ __forceinline__ double tutty_weight(int a, double delta, double c2)
{
#pragma clang fp contract(off)
    const double d1 = delta - 1.0;
    const double f = (1.0 - 2.0 * c2) - delta * delta;
    const double g = (1.0 - 2.0 * c2) - d1 * d1;
    switch (a) {
    case 0: return (1.0 - delta) * (0.5 * ((1.0 - f) - delta));
    case 1: return (1.0 - delta) * f + delta * (0.5 * ((1.0 - g) - d1));
    case 2: return (1.0 - delta) * (0.5 * ((1.0 - f) + delta)) + delta * g;
    default: return delta * (0.5 * ((1.0 - g) + d1));
    }
}

// (q, r) = (a / b, a % b) for 0 <= a, 0 < b, in 32-bit arithmetic when
// both fit (the per-system cell grids do), else 64-bit
__device__ __forceinline__ void divmod_cells(long long a, long long b, long long &q, long long &r)
{
    if (((unsigned long long)a | (unsigned long long)b) < (1ull << 32)) {
        const uint32_t q32 = (uint32_t)a / (uint32_t)b;
        q = q32;
        r = (long long)((uint32_t)a - q32 * (uint32_t)b);
    } else {
        q = a / b;
        r = a - q * b;
    }
}

}  // namespace phf
