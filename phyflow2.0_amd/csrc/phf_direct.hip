// phf_direct.hip -- all-pairs vortex-blob velocity sums on gfx950.
//
// Replaces the reference's direct kernels:
//   * vv2par_kernel (Gaussian, EPS-regularised),
//     src/cpp/blobs/cuda/vv2parGauss_gpu/vv2parGauss_device.cu:41-162
//   * direct_pot1_nomi_oseen_synchronized (fmm2d with tol == 0, Oseen),
//     src/cpp/blobs/c++/fmm2d_py_cpu_gpu/cudaevalkernels.h:61-71, 451-508.
//
// Mapping to CDNA4: one wave64 owns TPL*64 targets held in VGPRs; all
// lanes walk the same source sequence, so each source is fetched once per
// wave with scalar (SMEM) loads and broadcast through SGPR operands: no LDS
// staging and no barriers.  Long source lists are split over a second grid
// dimension; each (target block, source chunk) writes a float64 partial and
// a finishing pass adds the partials in a fixed order, so results do not
// depend on scheduling.
//
// Precision modes
//   F64 : float64 pair arithmetic (rcp + 2 Newton steps for 1/r2).
//   F32C: coordinates differenced in float64 and rounded once, pair
//         arithmetic in float32 (v_exp_f32, v_rcp_f32), sums in blocks of
//         16 terms folded into a TwoSum (hi, lo) accumulator per target.
#include "phf_ops.hpp"

#include <algorithm>
#include <cmath>

namespace phf {

// one source as a kernel uses it, read in place from the x, y, w arrays
struct alignas(32) Src64 {
    double x, y, w, pad;
};
struct alignas(8) Src32c {
    double x, y;
    float w, pad;
};

PairParams make_gauss_params(double ksigmasqr)
{
    PairParams p{};
    p.eps = 1e-12;
    p.inv_ks = 1.0 / ksigmasqr;
    p.c2 = 0.0;
    p.scale = 1.0;
    p.eps_f = (float)p.eps;
    p.nlog2e_ks_f = (float)(-1.4426950408889634074 / ksigmasqr);
    p.c2_f = 0.0f;
    p.scale_f = 1.0f;
    return p;
}

PairParams make_oseen_params(double xopt, double cutoff)
{
    // fmm.cpp:1166-1179 (pot=1, OSEEN, INLINECOMPLEX_DIRECT, cont=true)
    PairParams p{};
    p.eps = 0.0;
    p.c2 = cutoff * cutoff;
    p.inv_ks = 1.2564312086261696770 / (xopt * xopt);
    p.scale = -1.0 / std::expm1(-p.inv_ks * p.c2);
    p.eps_f = 0.0f;
    p.nlog2e_ks_f = (float)(-1.4426950408889634074 * p.inv_ks);
    p.c2_f = (float)p.c2;
    p.scale_f = (float)p.scale;
    p.f32_nocut = (p.scale_f == 1.0f && p.c2 * p.inv_ks * 1.4426950408889634074 > 25.5) ? 1 : 0;
    return p;
}

// ------------------------------------------------------ F32C kernel ------
template <int KIND, int TPL>
__global__ __launch_bounds__(64) void direct_f32c(
    const double *__restrict__ xb, const double *__restrict__ yb, const double *__restrict__ wb,
    int64_t nb, const double *__restrict__ xt,
    const double *__restrict__ yt, int64_t nt, int64_t chunk_len, PairParams pp,
    double2 *__restrict__ part)
{
    const int lane = threadIdx.x;
    const int64_t tbase = (int64_t)blockIdx.x * (kWave * TPL);
    double tx[TPL], ty[TPL];
    float hu[TPL], lu[TPL], hv[TPL], lv[TPL];
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
        int64_t t = tbase + lane + kWave * i;
        tx[i] = t < nt ? xt[t] : 0.0;
        ty[i] = t < nt ? yt[t] : 0.0;
        hu[i] = lu[i] = hv[i] = lv[i] = 0.0f;
    }
    const int64_t j0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t j1 = min(nb, j0 + chunk_len);
    int64_t j = j0;
    for (; j + 16 <= j1; j += 16) {
        float au[TPL], av[TPL];
#pragma unroll
        for (int i = 0; i < TPL; ++i) au[i] = av[i] = 0.0f;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const Src32c s{xb[j + k], yb[j + k], (float)wb[j + k], 0.0f};
#pragma unroll
            for (int i = 0; i < TPL; ++i) {
                float dx = (float)(tx[i] - s.x);
                float dy = (float)(ty[i] - s.y);
                float P = pair_f32<KIND>(dx, dy, s.w, pp);
                au[i] = __builtin_fmaf(P, dx, au[i]);
                av[i] = __builtin_fmaf(P, dy, av[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < TPL; ++i) {
            two_sum_add(hu[i], lu[i], au[i]);
            two_sum_add(hv[i], lv[i], av[i]);
        }
    }
    if (j < j1) {
        float au[TPL], av[TPL];
#pragma unroll
        for (int i = 0; i < TPL; ++i) au[i] = av[i] = 0.0f;
        for (; j < j1; ++j) {
            const Src32c s{xb[j], yb[j], (float)wb[j], 0.0f};
#pragma unroll
            for (int i = 0; i < TPL; ++i) {
                float dx = (float)(tx[i] - s.x);
                float dy = (float)(ty[i] - s.y);
                float P = pair_f32<KIND>(dx, dy, s.w, pp);
                au[i] = __builtin_fmaf(P, dx, au[i]);
                av[i] = __builtin_fmaf(P, dy, av[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < TPL; ++i) {
            two_sum_add(hu[i], lu[i], au[i]);
            two_sum_add(hv[i], lv[i], av[i]);
        }
    }
    double2 *out = part + (int64_t)blockIdx.y * nt;
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
        int64_t t = tbase + lane + kWave * i;
        if (t < nt)
            out[t] = make_double2((double)hu[i] + (double)lu[i], (double)hv[i] + (double)lv[i]);
    }
}

// ------------------------------------------------------- F64 kernel ------
template <int KIND, int TPL>
__global__ __launch_bounds__(64) void direct_f64(
    const double *__restrict__ xb, const double *__restrict__ yb, const double *__restrict__ wb,
    int64_t nb, const double *__restrict__ xt,
    const double *__restrict__ yt, int64_t nt, int64_t chunk_len, PairParams pp,
    double2 *__restrict__ part)
{
    const int lane = threadIdx.x;
    const int64_t tbase = (int64_t)blockIdx.x * (kWave * TPL);
    double tx[TPL], ty[TPL], au[TPL], av[TPL];
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
        int64_t t = tbase + lane + kWave * i;
        tx[i] = t < nt ? xt[t] : 0.0;
        ty[i] = t < nt ? yt[t] : 0.0;
        au[i] = av[i] = 0.0;
    }
    const int64_t j0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t j1 = min(nb, j0 + chunk_len);
    int64_t j = j0;
    for (; j + 8 <= j1; j += 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const Src64 s{xb[j + k], yb[j + k], wb[j + k], 0.0};
#pragma unroll
            for (int i = 0; i < TPL; ++i) {
                double dx = tx[i] - s.x;
                double dy = ty[i] - s.y;
                double P = pair_f64<KIND>(dx, dy, s.w, pp);
                au[i] = __builtin_fma(P, dx, au[i]);
                av[i] = __builtin_fma(P, dy, av[i]);
            }
        }
    }
    for (; j < j1; ++j) {
        const Src64 s{xb[j], yb[j], wb[j], 0.0};
#pragma unroll
        for (int i = 0; i < TPL; ++i) {
            double dx = tx[i] - s.x;
            double dy = ty[i] - s.y;
            double P = pair_f64<KIND>(dx, dy, s.w, pp);
            au[i] = __builtin_fma(P, dx, au[i]);
            av[i] = __builtin_fma(P, dy, av[i]);
        }
    }
    double2 *out = part + (int64_t)blockIdx.y * nt;
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
        int64_t t = tbase + lane + kWave * i;
        if (t < nt) out[t] = make_double2(au[i], av[i]);
    }
}

// Adds the chunk partials and maps (A, B) = (sum P dx, sum P dy) to the
// reference outputs:
//   Gaussian: vx = inv_2pi * (-B), vy = inv_2pi * A  (vv2parGauss_C_cpu.c:150-151)
//   cutoff:   the same                               (vv2par_C_cpu.c:149-150)
//   Oseen:    qr = -A,  qi = B                       (direct.cpp:1482-1483)
// A workgroup of 4 waves serves 64 targets: wave w adds the w-th quarter of
// the chunks in chunk order (coalesced rows, eight loads in flight), and
// the four quarter sums are added in wave order, so the result is fixed
// for a given chunking.
constexpr int FIN_T = 64;  // targets per finishing workgroup
__global__ __launch_bounds__(256) void direct_finish(const double2 *__restrict__ part,
                                                     int nchunk, int64_t nt, int kind,
                                                     double *__restrict__ o0,
                                                     double *__restrict__ o1)
{
    __shared__ double2 red[4][FIN_T];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * FIN_T + l;
    const int per = (nchunk + 3) / 4;
    const int c1 = min(nchunk, (w + 1) * per);
    double a = 0.0, b = 0.0;
    if (t < nt) {
        int c = w * per;
        for (; c + 8 <= c1; c += 8) {
            double2 p[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) p[u] = part[(int64_t)(c + u) * nt + t];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                a += p[u].x;
                b += p[u].y;
            }
        }
        for (; c < c1; ++c) {
            const double2 p = part[(int64_t)c * nt + t];
            a += p.x;
            b += p.y;
        }
    }
    red[w][l] = make_double2(a, b);
    __syncthreads();
    if (w != 0 || t >= nt) return;
    a = ((red[0][l].x + red[1][l].x) + red[2][l].x) + red[3][l].x;
    b = ((red[0][l].y + red[1][l].y) + red[2][l].y) + red[3][l].y;
    if (kind == PAIR_GAUSS_EPS || kind == PAIR_CUTOFF) {
        const double inv_2pi = 0.5 / 3.141592653589793;
        o0[t] = inv_2pi * (-b);
        o1[t] = inv_2pi * a;
    } else {
        o0[t] = -a;
        o1[t] = b;
    }
}

// ------------------------------------------------ vorticity (Gaussian) ----
// ww2parGauss_C_cpu.c:133-145:
//   w(t) = (1/pi) sum_m w_m exp(-(|t - z_m|^2 + EPS)/ks) / ks,  EPS = 1e-12
// Same layout as the velocity kernels with one accumulator per target.
template <int TPL, bool CUT>
__global__ __launch_bounds__(64) void vort_f32c(const double *__restrict__ xb,
                                                const double *__restrict__ yb,
                                                const double *__restrict__ wb, int64_t nb,
                                                const double *__restrict__ xt,
                                                const double *__restrict__ yt, int64_t nt,
                                                int64_t chunk_len, PairParams pp,
                                                double *__restrict__ part)
{
    const int lane = threadIdx.x;
    const int64_t tbase = (int64_t)blockIdx.x * (kWave * TPL);
    double tx[TPL], ty[TPL];
    float h[TPL], l[TPL];
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
        const int64_t t = tbase + lane + kWave * i;
        tx[i] = t < nt ? xt[t] : 0.0;
        ty[i] = t < nt ? yt[t] : 0.0;
        h[i] = l[i] = 0.0f;
    }
    const int64_t j0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t j1 = min(nb, j0 + chunk_len);
    for (int64_t j = j0; j < j1; j += 16) {
        const int64_t je = min(j1, j + 16);
        float a[TPL];
#pragma unroll
        for (int i = 0; i < TPL; ++i) a[i] = 0.0f;
        for (int64_t k = j; k < je; ++k) {
            const Src32c s{xb[k], yb[k], (float)wb[k], 0.0f};
#pragma unroll
            for (int i = 0; i < TPL; ++i) {
                const float dx = (float)(tx[i] - s.x), dy = (float)(ty[i] - s.y);
                const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, pp.eps_f));
                const float f = CUT ? __builtin_amdgcn_rsqf(r2) : exp2_f32(r2 * pp.nlog2e_ks_f);
                a[i] = __builtin_fmaf(s.w, f, a[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < TPL; ++i) two_sum_add(h[i], l[i], a[i]);
    }
    double *out = part + (int64_t)blockIdx.y * nt;
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
        const int64_t t = tbase + lane + kWave * i;
        if (t < nt) out[t] = (double)h[i] + (double)l[i];
    }
}

template <int TPL, bool CUT>
__global__ __launch_bounds__(64) void vort_f64(const double *__restrict__ xb,
                                               const double *__restrict__ yb,
                                               const double *__restrict__ wb, int64_t nb,
                                               const double *__restrict__ xt,
                                               const double *__restrict__ yt, int64_t nt,
                                               int64_t chunk_len, PairParams pp,
                                               double *__restrict__ part)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x;
    const int64_t tbase = (int64_t)blockIdx.x * (kWave * TPL);
    double tx[TPL], ty[TPL], a[TPL];
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
        const int64_t t = tbase + lane + kWave * i;
        tx[i] = t < nt ? xt[t] : 0.0;
        ty[i] = t < nt ? yt[t] : 0.0;
        a[i] = 0.0;
    }
    const int64_t j0 = (int64_t)blockIdx.y * chunk_len;
    const int64_t j1 = min(nb, j0 + chunk_len);
    for (int64_t j = j0; j < j1; ++j) {
        const Src64 s{xb[j], yb[j], wb[j], 0.0};
#pragma unroll
        for (int i = 0; i < TPL; ++i) {
            const double dx = tx[i] - s.x, dy = ty[i] - s.y;
            const double r2 = dx * dx + dy * dy + pp.eps;
            a[i] += CUT ? s.w / sqrt(r2) : s.w * exp(-r2 * pp.inv_ks);
        }
    }
    double *out = part + (int64_t)blockIdx.y * nt;
#pragma unroll
    for (int i = 0; i < TPL; ++i) {
        const int64_t t = tbase + lane + kWave * i;
        if (t < nt) out[t] = a[i];
    }
}

// chunk partials in chunk order, then inv_pi * w / ks (ww2parGauss_C_cpu.c:144)
// or, for the cutoff kernel, inv_2pi * w (ww2par_C_cpu.c:144)
__global__ void vort_finish(const double *__restrict__ part, int nchunk, int64_t nt, double ks,
                            int cut, double *__restrict__ w)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    double a = 0.0;
    int c = 0;
    for (; c + 8 <= nchunk; c += 8) {
        double p[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) p[u] = part[(int64_t)(c + u) * nt + t];
#pragma unroll
        for (int u = 0; u < 8; ++u) a += p[u];
    }
    for (; c < nchunk; ++c) a += part[(int64_t)c * nt + t];
    if (cut) {
        const double inv_2pi = 0.5 / 3.141592653589793;
        w[t] = inv_2pi * a;
    } else {
        const double inv_pi = 1.0 / 3.141592653589793;
        w[t] = inv_pi * a / ks;
    }
}

// ------------------------------------------------------------- host ------
int upload(Lane *L, const double *host, int64_t n, double **dev)
{
    double *d;
    PHF_ALLOC(d, L->arena, double, (size_t)(n > 0 ? n : 1));
    if (n > 0)
        PHF_HIP(hipMemcpyAsync(d, host, sizeof(double) * n, hipMemcpyHostToDevice, L->stream));
    *dev = d;
    return PHF_OK;
}

int download(Lane *L, const double *dev, int64_t n, double *host)
{
    if (n > 0)
        PHF_HIP(hipMemcpyAsync(host, dev, sizeof(double) * n, hipMemcpyDeviceToHost, L->stream));
    return PHF_OK;
}

// Source chunks of a direct sum over nt targets in groups of 64 * tpl:
// enough chunks that the grid holds ~8 waves per SIMD of the chip, each of
// at least 64 sources (a multiple of 16); one chunk once the target groups
// alone fill the chip.  max_chunks caps the count (the partial sums take
// nchunk * nt double2).
static void direct_split(int64_t nt, int tpl, int64_t nb, int64_t *nchunk_out,
                         int64_t *chunk_len_out, int64_t max_chunks = 65535)
{
    const int64_t tblock = (int64_t)kWave * tpl;
    const int64_t ngroups = (nt + tblock - 1) / tblock;
    const int64_t want_waves = 8192;
    int64_t nchunk = std::max<int64_t>(1, (want_waves + ngroups - 1) / ngroups);
    nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(1, nb / 64));
    nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(1, std::min<int64_t>(max_chunks, 65535)));
    int64_t chunk_len = (nb + nchunk - 1) / nchunk;
    chunk_len = std::max<int64_t>(16, (chunk_len + 15) / 16 * 16);
    *nchunk_out = std::max<int64_t>(1, (nb + chunk_len - 1) / chunk_len);
    *chunk_len_out = chunk_len;
}

// Zero partial row for an empty source set (a kernel rather than a memset,
// so the sequence holds only kernel launches).
__global__ __launch_bounds__(256) void zero_part(double2 *__restrict__ part, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) part[i] = make_double2(0.0, 0.0);
}

constexpr int DIRECT_TPL32 = 2, DIRECT_TPL64 = 1;

// Enqueue the pair kernel and the finishing kernel on stream st with the
// partial sums in part (nchunk * nt double2; nchunk, chunk_len from
// direct_split).
template <int KIND>
static int enqueue_direct(hipStream_t st, int prec, const PairParams &pp, const double *xb,
                          const double *yb, const double *wb, int64_t nb, const double *xt,
                          const double *yt, int64_t nt, int64_t nchunk, int64_t chunk_len,
                          double2 *part, double *o0, double *o1)
{
    const int tpl = prec == PHF_PREC_F32C ? DIRECT_TPL32 : DIRECT_TPL64;
    const int64_t tblock = (int64_t)kWave * tpl;
    const int64_t ngroups = (nt + tblock - 1) / tblock;
    dim3 grid((unsigned)ngroups, (unsigned)nchunk);
    if (nb == 0) {
        zero_part<<<div_up(nt, 256), 256, 0, st>>>(part, nt);
    } else if (prec == PHF_PREC_F32C) {
        // sources read in place: every lane of a wave reads the same one
        direct_f32c<KIND, DIRECT_TPL32><<<grid, kWave, 0, st>>>(xb, yb, wb, nb, xt, yt, nt,
                                                                chunk_len, pp, part);
    } else {
        direct_f64<KIND, DIRECT_TPL64><<<grid, kWave, 0, st>>>(xb, yb, wb, nb, xt, yt, nt,
                                                               chunk_len, pp, part);
    }
    PHF_HIP(hipGetLastError());
    direct_finish<<<div_up(nt, FIN_T), 256, 0, st>>>(
        part, nb == 0 ? 1 : (int)nchunk, nt,
        (KIND == PAIR_GAUSS_EPS || KIND == PAIR_CUTOFF) ? KIND : PAIR_OSEEN, o0, o1);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

template <int KIND>
static int launch_direct(Lane *L, int prec, const PairParams &pp, const double *xb,
                         const double *yb, const double *wb, int64_t nb, const double *xt,
                         const double *yt, int64_t nt, double *o0, double *o1)
{
    const int tpl = prec == PHF_PREC_F32C ? DIRECT_TPL32 : DIRECT_TPL64;
    int64_t nchunk, chunk_len;
    direct_split(nt, tpl, nb, &nchunk, &chunk_len);
    double2 *part;
    PHF_ALLOC(part, L->arena, double2, (size_t)nchunk * (size_t)nt);
    return enqueue_direct<KIND>(L->stream, prec, pp, xb, yb, wb, nb, xt, yt, nt, nchunk,
                                chunk_len, part, o0, o1);
}

int vorticity_device(Lane *L, int prec, double ksigmasqr, const double *xb, const double *yb,
                     const double *wb, int64_t nb, const double *xt, const double *yt, int64_t nt,
                     double *w, bool cutoff)
{
    if (nt == 0) return PHF_OK;
    PairParams pp = make_gauss_params(ksigmasqr);
    if (cutoff) {  // sigma^2 takes the place of EPS (ww2par_C_cpu.c:140)
        pp.eps = ksigmasqr;
        pp.eps_f = (float)ksigmasqr;
    }
    constexpr int TPL32 = 2, TPL64 = 1;
    const int tpl = prec == PHF_PREC_F32C ? TPL32 : TPL64;
    const int64_t tblock = (int64_t)kWave * tpl;
    const int64_t ngroups = (nt + tblock - 1) / tblock;
    int64_t nchunk, chunk_len;
    direct_split(nt, tpl, nb, &nchunk, &chunk_len);
    double *part;
    PHF_ALLOC(part, L->arena, double, (size_t)nchunk * (size_t)nt);
    const dim3 grid((unsigned)ngroups, (unsigned)nchunk);
    if (nb == 0) {
        PHF_HIP(hipMemsetAsync(part, 0, sizeof(double) * nt, L->stream));
    } else if (prec == PHF_PREC_F32C) {
        if (cutoff)
            vort_f32c<TPL32, true><<<grid, kWave, 0, L->stream>>>(xb, yb, wb, nb, xt, yt, nt,
                                                                  chunk_len, pp, part);
        else
            vort_f32c<TPL32, false><<<grid, kWave, 0, L->stream>>>(xb, yb, wb, nb, xt, yt, nt,
                                                                   chunk_len, pp, part);
    } else {
        if (cutoff)
            vort_f64<TPL64, true><<<grid, kWave, 0, L->stream>>>(xb, yb, wb, nb, xt, yt, nt,
                                                                 chunk_len, pp, part);
        else
            vort_f64<TPL64, false><<<grid, kWave, 0, L->stream>>>(xb, yb, wb, nb, xt, yt, nt,
                                                                  chunk_len, pp, part);
    }
    PHF_HIP(hipGetLastError());
    vort_finish<<<div_up(nt, 256), 256, 0, L->stream>>>(part, nb == 0 ? 1 : (int)nchunk, nt,
                                                         ksigmasqr, cutoff ? 1 : 0, w);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

int direct_device(Lane *L, int kind, int prec, const PairParams &pp, const double *xb,
                  const double *yb, const double *wb, int64_t nb, const double *xt,
                  const double *yt, int64_t nt, double *o0, double *o1)
{
    if (nt == 0) return PHF_OK;
    if (kind == PAIR_GAUSS_EPS)
        return launch_direct<PAIR_GAUSS_EPS>(L, prec, pp, xb, yb, wb, nb, xt, yt, nt, o0, o1);
    if (kind == PAIR_CUTOFF)
        return launch_direct<PAIR_CUTOFF>(L, prec, pp, xb, yb, wb, nb, xt, yt, nt, o0, o1);
    if (prec == PHF_PREC_F32C && pp.f32_nocut)
        return launch_direct<PAIR_OSEEN_NOCUT>(L, prec, pp, xb, yb, wb, nb, xt, yt, nt, o0, o1);
    return launch_direct<PAIR_OSEEN>(L, prec, pp, xb, yb, wb, nb, xt, yt, nt, o0, o1);
}

}  // namespace phf

// ------------------------------------------------------------ C entry ----
using namespace phf;

static int vv2par_entry(int kind, const double *xb, const double *yb, const double *wb,
                        int64_t nb, const double *xt, const double *yt, int64_t nt,
                        double ksigmasqr, int32_t prec, double *vx, double *vy);

extern "C" int phf_vv2par_gauss(const double *xb, const double *yb, const double *wb,
                                int64_t nb, const double *xt, const double *yt, int64_t nt,
                                double ksigmasqr, int32_t prec, double *vx, double *vy)
{
    clear_error();
    return vv2par_entry(PAIR_GAUSS_EPS, xb, yb, wb, nb, xt, yt, nt, ksigmasqr, prec, vx, vy);
}

extern "C" int phf_vv2par_cutoff(const double *xb, const double *yb, const double *wb,
                                 int64_t nb, const double *xt, const double *yt, int64_t nt,
                                 double sigmasqr, int32_t prec, double *vx, double *vy)
{
    clear_error();
    return vv2par_entry(PAIR_CUTOFF, xb, yb, wb, nb, xt, yt, nt, sigmasqr, prec, vx, vy);
}

static int vv2par_entry(int kind, const double *xb, const double *yb, const double *wb,
                        int64_t nb, const double *xt, const double *yt, int64_t nt,
                        double ksigmasqr, int32_t prec, double *vx, double *vy)
{
    PHF_REQUIRE(nb >= 0 && nt >= 0, "negative sizes (nb=%lld, nt=%lld)", (long long)nb,
                (long long)nt);
    PHF_REQUIRE(prec == PHF_PREC_F64 || prec == PHF_PREC_F32C, "unknown precision %d", prec);
    PHF_REQUIRE(std::isfinite(ksigmasqr) && ksigmasqr > 0.0,
                "ksigmasqr must be finite and > 0 (got %g)", ksigmasqr);
    PHF_TRY(check_host_array(xb, nb, "xBlob"));
    PHF_TRY(check_host_array(yb, nb, "yBlob"));
    PHF_TRY(check_host_array(wb, nb, "wBlob"));
    PHF_TRY(check_host_array(xt, nt, "xTarget"));
    PHF_TRY(check_host_array(yt, nt, "yTarget"));
    PHF_REQUIRE(nt == 0 || (vx && vy), "output arrays are NULL");
    if (nt == 0) return PHF_OK;

    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *L = g.lane;
    PHF_TRY(reset_arena(L));
    StageTimer tm(L, timing_enabled());
    bool used[PHF_STAGE_COUNT] = {};
    used[PHF_STAGE_UPLOAD] = used[PHF_STAGE_P2P] = used[PHF_STAGE_DOWNLOAD] = true;
    used[PHF_STAGE_TOTAL] = true;

    tm.begin(PHF_STAGE_TOTAL);
    tm.begin(PHF_STAGE_UPLOAD);
    double *dxb, *dyb, *dwb, *dxt, *dyt, *o0, *o1;
    PHF_TRY(upload(L, xb, nb, &dxb));
    PHF_TRY(upload(L, yb, nb, &dyb));
    PHF_TRY(upload(L, wb, nb, &dwb));
    PHF_TRY(upload(L, xt, nt, &dxt));
    PHF_TRY(upload(L, yt, nt, &dyt));
    PHF_ALLOC(o0, L->arena, double, (size_t)nt);
    PHF_ALLOC(o1, L->arena, double, (size_t)nt);
    tm.end(PHF_STAGE_UPLOAD);
    tm.begin(PHF_STAGE_P2P);
    PairParams pp = make_gauss_params(ksigmasqr);
    if (kind == PAIR_CUTOFF) {  // r2 = dx^2 + dy^2 + sigma^2 (vv2par_C_cpu.c:144)
        pp.eps = ksigmasqr;
        pp.eps_f = (float)ksigmasqr;
    }
    PHF_TRY(direct_device(L, kind, prec, pp, dxb, dyb, dwb, nb, dxt, dyt, nt, o0, o1));
    tm.end(PHF_STAGE_P2P);
    tm.begin(PHF_STAGE_DOWNLOAD);
    PHF_TRY(download(L, o0, nt, vx));
    PHF_TRY(download(L, o1, nt, vy));
    tm.end(PHF_STAGE_DOWNLOAD);
    tm.end(PHF_STAGE_TOTAL);
    PHF_HIP(hipStreamSynchronize(L->stream));
    record_stage_times(L, used);
    return PHF_OK;
}

static int ww2par_entry(bool cutoff, const double *xb, const double *yb, const double *wb,
                        int64_t nb, const double *xt, const double *yt, int64_t nt,
                        double ksigmasqr, int32_t prec, double *w);

extern "C" int phf_ww2par_gauss(const double *xb, const double *yb, const double *wb,
                                int64_t nb, const double *xt, const double *yt, int64_t nt,
                                double ksigmasqr, int32_t prec, double *w)
{
    clear_error();
    return ww2par_entry(false, xb, yb, wb, nb, xt, yt, nt, ksigmasqr, prec, w);
}

extern "C" int phf_ww2par_cutoff(const double *xb, const double *yb, const double *wb,
                                 int64_t nb, const double *xt, const double *yt, int64_t nt,
                                 double sigmasqr, int32_t prec, double *w)
{
    clear_error();
    return ww2par_entry(true, xb, yb, wb, nb, xt, yt, nt, sigmasqr, prec, w);
}

static int ww2par_entry(bool cutoff, const double *xb, const double *yb, const double *wb,
                        int64_t nb, const double *xt, const double *yt, int64_t nt,
                        double ksigmasqr, int32_t prec, double *w)
{
    PHF_REQUIRE(nb >= 0 && nt >= 0, "negative sizes (nb=%lld, nt=%lld)", (long long)nb,
                (long long)nt);
    PHF_REQUIRE(prec == PHF_PREC_F64 || prec == PHF_PREC_F32C, "unknown precision %d", prec);
    PHF_REQUIRE(std::isfinite(ksigmasqr) && ksigmasqr > 0.0,
                "ksigmasqr must be finite and > 0 (got %g)", ksigmasqr);
    PHF_TRY(check_host_array(xb, nb, "xBlob"));
    PHF_TRY(check_host_array(yb, nb, "yBlob"));
    PHF_TRY(check_host_array(wb, nb, "wBlob"));
    PHF_TRY(check_host_array(xt, nt, "xTarget"));
    PHF_TRY(check_host_array(yt, nt, "yTarget"));
    PHF_REQUIRE(nt == 0 || w, "output array is NULL");
    if (nt == 0) return PHF_OK;
    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *L = g.lane;
    PHF_TRY(reset_arena(L));
    double *dxb, *dyb, *dwb, *dxt, *dyt, *o;
    PHF_TRY(upload(L, xb, nb, &dxb));
    PHF_TRY(upload(L, yb, nb, &dyb));
    PHF_TRY(upload(L, wb, nb, &dwb));
    PHF_TRY(upload(L, xt, nt, &dxt));
    PHF_TRY(upload(L, yt, nt, &dyt));
    PHF_ALLOC(o, L->arena, double, (size_t)nt);
    PHF_TRY(vorticity_device(L, prec, ksigmasqr, dxb, dyb, dwb, nb, dxt, dyt, nt, o, cutoff));
    PHF_TRY(download(L, o, nt, w));
    PHF_HIP(hipStreamSynchronize(L->stream));
    return PHF_OK;
}

// ------------------------------------------- allocation-free direct sum ---
// phf_vv2par_gauss_noalloc keeps its partial sums in one file-scope device
// array instead of a workspace from hipMalloc.  Each GPU the library runs
// on holds its own copy until the process ends.  Host code never uses the
// array's name as an address: direct_scratch_setup (called by the library's
// init with the device current) asks the runtime for the device address
// and every launch takes that pointer.
namespace phf {

constexpr int64_t kScratchPart = int64_t(1) << 21;  // double2 entries: 32 MiB
constexpr int kScratchDevices = 64;
static __device__ __align__(256) double2 scratch[kScratchPart];
static double2 *g_scratch[kScratchDevices] = {};

int direct_scratch_setup(int dev)
{
    PHF_REQUIRE(dev >= 0 && dev < kScratchDevices, "device %d out of range", dev);
    if (g_scratch[dev]) return PHF_OK;
    void *p = nullptr;
    PHF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PHF_REQUIRE(p != nullptr, "no device address for the direct-sum scratch array");
    g_scratch[dev] = reinterpret_cast<double2 *>(p);
    return PHF_OK;
}

// A kernel argument must lie inside one device (or managed) allocation of
// the current device, so a host pointer fails here instead of faulting the
// GPU.
static int check_device_range(const void *p, int64_t n, const char *name)
{
    if (n == 0) return PHF_OK;
    PHF_REQUIRE(p != nullptr, "%s is NULL but has %lld entries", name, (long long)n);
    hipPointerAttribute_t at{};
    hipError_t e = hipPointerGetAttributes(&at, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("%s is not device memory (%s)", name, hipGetErrorName(e));
        return PHF_ERR_INVALID;
    }
    PHF_REQUIRE(at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged,
                "%s is not device memory (memory type %d)", name, (int)at.type);
    hipDeviceptr_t base = nullptr;
    size_t bytes = 0;
    e = hipMemGetAddressRange(&base, &bytes, const_cast<void *>(p));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("%s: allocation range unknown (%s)", name, hipGetErrorName(e));
        return PHF_ERR_INVALID;
    }
    const char *b = reinterpret_cast<const char *>(base);
    const char *q = reinterpret_cast<const char *>(p);
    PHF_REQUIRE(q >= b && (uint64_t)(q - b) + (uint64_t)n * sizeof(double) <= bytes,
                "%s: %lld doubles run past the end of its allocation", name, (long long)n);
    return PHF_OK;
}

}  // namespace phf

extern "C" int phf_vv2par_gauss_noalloc(const double *xb, const double *yb, const double *wb,
                                        int64_t nb, const double *xt, const double *yt,
                                        int64_t nt, double ksigmasqr, int32_t prec, double *vx,
                                        double *vy, void *stream)
{
    clear_error();
    PHF_REQUIRE(nb >= 0 && nt >= 0, "negative sizes (nb=%lld, nt=%lld)", (long long)nb,
                (long long)nt);
    PHF_REQUIRE(nt <= kScratchPart,
                "nt=%lld exceeds the %lld targets the allocation-free scratch holds",
                (long long)nt, (long long)kScratchPart);
    PHF_REQUIRE(nb < (int64_t(1) << 40), "nb out of range (%lld)", (long long)nb);
    PHF_REQUIRE(prec == PHF_PREC_F64 || prec == PHF_PREC_F32C, "unknown precision %d", prec);
    PHF_REQUIRE(std::isfinite(ksigmasqr) && ksigmasqr > 0.0,
                "ksigmasqr must be finite and > 0 (got %g)", ksigmasqr);
    if (nt == 0) return PHF_OK;
    PHF_TRY(Context::get().init());
    const int dev = Context::get().device();
    PHF_HIP(hipSetDevice(dev));
    PHF_REQUIRE(dev >= 0 && dev < kScratchDevices && g_scratch[dev] != nullptr,
                "the direct-sum scratch array of device %d is not set up", dev);
    PHF_TRY(check_device_range(xb, nb, "xBlob"));
    PHF_TRY(check_device_range(yb, nb, "yBlob"));
    PHF_TRY(check_device_range(wb, nb, "wBlob"));
    PHF_TRY(check_device_range(xt, nt, "xTarget"));
    PHF_TRY(check_device_range(yt, nt, "yTarget"));
    PHF_TRY(check_device_range(vx, nt, "vx"));
    PHF_TRY(check_device_range(vy, nt, "vy"));

    const int tpl = prec == PHF_PREC_F32C ? DIRECT_TPL32 : DIRECT_TPL64;
    int64_t nchunk, chunk_len;
    direct_split(nt, tpl, nb, &nchunk, &chunk_len, kScratchPart / nt);
    PHF_REQUIRE(nchunk * nt <= kScratchPart, "partial sums (%lld x %lld) exceed the scratch",
                (long long)nchunk, (long long)nt);
    const PairParams pp = make_gauss_params(ksigmasqr);
    if (stream) {
        // enqueue on the caller's stream; the caller synchronises
        return enqueue_direct<PAIR_GAUSS_EPS>(reinterpret_cast<hipStream_t>(stream), prec, pp,
                                              xb, yb, wb, nb, xt, yt, nt, nchunk, chunk_len,
                                              g_scratch[dev], vx, vy);
    }
    LaneGuard g;
    PHF_TRY(g.rc);
    PHF_TRY(enqueue_direct<PAIR_GAUSS_EPS>(g.lane->stream, prec, pp, xb, yb, wb, nb, xt, yt, nt,
                                           nchunk, chunk_len, g_scratch[dev], vx, vy));
    PHF_HIP(hipStreamSynchronize(g.lane->stream));
    return PHF_OK;
}
