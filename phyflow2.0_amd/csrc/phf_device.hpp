// phf_device.hpp -- device-side helpers shared by the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace phf {

constexpr int kWave = 64;

// Pair-kernel selector.
enum PairKind : int {
    PAIR_GAUSS_EPS = 0,  // vv2parGauss: r2 += 1e-12, s = w/r2*(1-exp(-r2/ks))
    PAIR_OSEEN = 1,      // fmm2d: P = w/r2 * (r2 < c2 ? -expm1(-shape r2)*scale : 1)
    // Oseen in float32 when the cutoff branch is exact without the select:
    // scale rounds to 1.0f and exp(-shape c2) < 2^-25, so 1 - e == 1.0f for
    // every r2 >= c2 (the default cutoff 5 xopt gives exp(-31.4)).
    PAIR_OSEEN_NOCUT = 2,
    // algebraic cutoff kernel (vv2par_C_cpu.c:142-147): P = w/(r2 + sigma^2),
    // sigma^2 carried in eps
    PAIR_CUTOFF = 3,
};

// Parameters of the pair kernel, prepared on the host in float64.
struct PairParams {
    double eps;       // added to r2 (Gaussian: EPS; cutoff: sigma^2)
    double inv_ks;    // 1/(k sigma^2) == shape of the Oseen kernel
    double c2;        // cutoff^2 (Oseen)
    double scale;     // -1/expm1(-shape c2) (Oseen)
    float eps_f, nlog2e_ks_f, c2_f, scale_f;  // float32 forms
    int f32_nocut;    // Oseen cutoff branch is a no-op in float32
};

// Error-free transformation: (hi, lo) += x, Knuth's TwoSum.
__device__ __forceinline__ void two_sum_add(float &hi, float &lo, float x)
{
    float s = hi + x;
    float bp = s - hi;
    float err = (hi - (s - bp)) + (x - bp);
    hi = s;
    lo += err;
}

// float64 reciprocal: hardware estimate refined by two Newton steps.  On
// gfx950 the estimate (v_rcp_f64) is good to 2^-24.4 relative, one step
// leaves up to 20 ulp and two gave the correctly rounded 1/x for 4M inputs
// spread over 2^-80 .. 2^80 (max relative error against IEEE division).
__device__ __forceinline__ double rcp_f64(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    return y;
}

__device__ __forceinline__ float rcp_f32(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float exp2_f32(float x) { return __builtin_amdgcn_exp2f(x); }

// Interleave the low 16 bits of v onto the even bits.
__host__ __device__ __forceinline__ uint32_t spread_bits(uint32_t v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__host__ __device__ __forceinline__ uint32_t compact_bits(uint32_t v)
{
    v &= 0x55555555u;
    v = (v | (v >> 1)) & 0x33333333u;
    v = (v | (v >> 2)) & 0x0F0F0F0Fu;
    v = (v | (v >> 4)) & 0x00FF00FFu;
    v = (v | (v >> 8)) & 0x0000FFFFu;
    return v;
}

__host__ __device__ __forceinline__ uint32_t morton2(uint32_t ix, uint32_t iy)
{
    return spread_bits(ix) | (spread_bits(iy) << 1);
}

// ------------------------------------------------------------ pair ops --
// P for one (target, source) pair; the caller accumulates P*dx and P*dy
// with dx = x_target - x_source.
//   Gaussian (vv2parGauss_C_cpu.c:144-146): P = w/(r2+eps) * (1 - exp(-(r2+eps)/ks))
//   Oseen (direct.cpp:1428-1485):           P = w/r2 * (r2 < c2 ? -expm1(-r2/ks)*scale : 1),
//                                            P = 0 for r2 == 0.
// float32: the Oseen r2 carries +1e-30 so a coincident pair gives
// 1 - exp2(-tiny) = 0 exactly and P = 0 without a select; w (1 - e) is one
// FMA.  Per pair: 9 VALU + 2 transcendental instructions.
template <int KIND>
__device__ __forceinline__ float pair_f32(float dx, float dy, float w, const PairParams &pp)
{
    if (KIND == PAIR_GAUSS_EPS) {
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, pp.eps_f));
        const float e = exp2_f32(r2 * pp.nlog2e_ks_f);
        return __builtin_fmaf(-w, e, w) * rcp_f32(r2);
    } else if (KIND == PAIR_CUTOFF) {
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, pp.eps_f));
        return w * rcp_f32(r2);
    } else if (KIND == PAIR_OSEEN_NOCUT) {
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, 1e-30f));
        const float e = exp2_f32(r2 * pp.nlog2e_ks_f);
        return __builtin_fmaf(-w, e, w) * rcp_f32(r2);
    } else {
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, 1e-30f));
        const float e = exp2_f32(r2 * pp.nlog2e_ks_f);
        const float ws = w * pp.scale_f;
        const float num = r2 < pp.c2_f ? __builtin_fmaf(-ws, e, ws) : w;
        return num * rcp_f32(r2);
    }
}

// pair_f32 for Oseen pairs known to satisfy exp(-shape r2) <= 2^-26: then
// w (1 - e) rounds to w in float32 (|w e| is below half an ulp of w), so
// dropping the exponential gives bit-identical P.  7 VALU + 1 transcendental.
template <int KIND>
__device__ __forceinline__ float pair_f32_far(float dx, float dy, float w, const PairParams &pp)
{
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, 1e-30f));
    if (KIND == PAIR_OSEEN) {
        const float num = r2 < pp.c2_f ? w * pp.scale_f : w;
        return num * rcp_f32(r2);
    }
    return w * rcp_f32(r2);
}

// expm1(x) for x <= 0 in float64, within 1.5 ulp (checked against a
// long-double expm1 over [-50, 0] including subnormal x), with no branch:
// Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 = LN2_HI +
// LN2_LO and LN2_HI carrying 32 significant bits so that k LN2_HI is exact;
// expm1(r) by its Taylor series through r^13 (truncation below 2e-17
// relative at |r| = ln2/2); expm1(x) = 2^k expm1(r) + (2^k - 1), whose
// second term is exact for k >= -53.  x is clamped at -40, where expm1
// rounds to -1.
// a b + c as the three-operand v_fma_f64.  With c a constant held in a
// register (a Horner coefficient hoisted out of a loop) the compiler
// otherwise copies c into the destination and uses the two-address
// v_fmac_f64: one extra move per Horner step.  The same operation, so the
// same result.
__device__ __forceinline__ double fma3_f64(double a, double b, double c)
{
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
    return d;
}

__device__ __forceinline__ double expm1_neg_f64(double x)
{
    x = __builtin_fmax(x, -40.0);
    const double k = __builtin_rint(x * 1.4426950408889634);
    double r = __builtin_fma(-k, 0x1.62e42ff000000p-1, x);
    r = __builtin_fma(-k, -0x1.718432a1b0e26p-35, r);
    // q = sum_{j=0..12} r^j / (j+1)!
    double q = 1.6059043836821613e-10;
    q = fma3_f64(q, r, 2.08767569878681e-09);
    q = fma3_f64(q, r, 2.505210838544172e-08);
    q = fma3_f64(q, r, 2.7557319223985888e-07);
    q = fma3_f64(q, r, 2.7557319223985893e-06);
    q = fma3_f64(q, r, 2.4801587301587302e-05);
    q = fma3_f64(q, r, 1.9841269841269841e-04);
    q = fma3_f64(q, r, 1.3888888888888889e-03);
    q = fma3_f64(q, r, 8.3333333333333332e-03);
    q = fma3_f64(q, r, 4.1666666666666664e-02);
    q = fma3_f64(q, r, 1.6666666666666666e-01);
    q = fma3_f64(q, r, 0.5);
    q = fma3_f64(q, r, 1.0);
    const double t = __builtin_amdgcn_ldexp(1.0, (int)k);
    return __builtin_fma(t, r * q, t - 1.0);
}

template <int KIND>
__device__ __forceinline__ double pair_f64(double dx, double dy, double w, const PairParams &pp)
{
    if (KIND == PAIR_GAUSS_EPS) {
        // 1 - exp(-r2/ks) as -expm1 (agrees with the subtraction to an ulp
        // of 1, without its cancellation at small r2)
        const double r2 = __builtin_fma(dx, dx, dy * dy) + pp.eps;
        return (w * rcp_f64(r2)) * -expm1_neg_f64(-r2 * pp.inv_ks);
    } else if (KIND == PAIR_CUTOFF) {
        const double r2 = __builtin_fma(dx, dx, dy * dy) + pp.eps;
        return w * rcp_f64(r2);
    } else {
        // scale (1 - e) rises through 1 exactly at r2 = c2 (scale =
        // 1/(1 - exp(-shape c2))), so the cutoff branch is a min; with
        // cutoff 0 the scale is +inf and every r2 > 0 gives 1
        const double r2 = __builtin_fma(dx, dx, dy * dy);
        const double om = __builtin_fmin(-expm1_neg_f64(-pp.inv_ks * r2) * pp.scale, 1.0);
        const double P = (w * rcp_f64(r2)) * om;
        return r2 > 0.0 ? P : 0.0;
    }
}

}  // namespace phf
