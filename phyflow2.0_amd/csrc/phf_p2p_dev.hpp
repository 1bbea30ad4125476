// phf_p2p_dev.hpp -- the float32 near-field sum over LDS-staged sources,
// shared by the single-system and the batched FMM.
//
// A workgroup of 4 waves serves one target leaf.  The sources of the
// leaf's near leaves (ranges padded to multiples of 4 with zero-weight
// entries) are staged in LDS in batches, shifted into the target leaf's
// frame; every wave sweeps a quarter of each batch for every target of the
// leaf (p2p_leaf), and the per-wave partials meet in LDS.
#pragma once

#include "phf_device.hpp"

namespace phf {

constexpr int P2P_STAGE = 1536;  // staged sources (float4) per batch

// Pair op in scaled coordinates x' = lam x with lam^2 = shape log2(e): the
// smoothing factor is exp2(-r2') (the negation is a free source modifier),
// one multiply per pair fewer than pair_f32; sums come out divided by lam.
// CORE = false: exp(-shape r2) <= 2^-26 for the pair, so w (1 - e) rounds to
// w and the exponential is dropped (as pair_f32_far).
template <int KIND, bool CORE>
__device__ __forceinline__ float pair_scaled(float dx, float dy, float w, float c2s,
                                             float scale_f)
{
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, 1e-30f));
    if (CORE) {
        const float e = exp2_f32(-r2);
        if (KIND == PAIR_OSEEN) {
            const float ws = w * scale_f;
            const float num = r2 < c2s ? __builtin_fmaf(-ws, e, ws) : w;
            return num * rcp_f32(r2);
        }
        return __builtin_fmaf(-w, e, w) * rcp_f32(r2);
    }
    if (KIND == PAIR_OSEEN) {
        const float num = r2 < c2s ? w * scale_f : w;
        return num * rcp_f32(r2);
    }
    return w * rcp_f32(r2);
}

__device__ __forceinline__ void two_sum_merge(float &h, float &l, float h2, float l2)
{
    const float s = h + h2;
    const float bp = s - h;
    const float err = (h - (s - bp)) + (h2 - bp);
    h = s;
    l = (l + l2) + err;
}


// every lane of a target ends with the sum of its four groups: (0,1) and
// (2,3), then the pairs (a fixed butterfly)
__device__ __forceinline__ void p2p_merge_groups(float &ha, float &la, float &hb, float &lb)
{
    two_sum_merge(ha, la, __shfl_xor(ha, 16), __shfl_xor(la, 16));
    two_sum_merge(hb, lb, __shfl_xor(hb, 16), __shfl_xor(lb, 16));
    two_sum_merge(ha, la, __shfl_xor(ha, 32), __shfl_xor(la, 32));
    two_sum_merge(hb, lb, __shfl_xor(hb, 32), __shfl_xor(lb, 32));
}

// local expansion sum_k B_k z^k (Horner, float32)
__device__ __forceinline__ void l2p_f32(const float2 *__restrict__ Bb, int p1, float zr, float zi,
                                        double &fr, double &fi)
{
    float pr = 0.f, pi = 0.f;
    for (int k = p1 - 1; k >= 0; --k) {
        const float2 c = Bb[k];
        const float nr = pr * zr - pi * zi + c.x;
        const float ni = pr * zi + pi * zr + c.y;
        pr = nr;
        pi = ni;
    }
    fr = pr;
    fi = pi;
}

// Staging plan of one batch: lane l of every wave looks at near index q + l,
// a wave scan of the range lengths packs as many as fit in STAGE entries
// (the cursor (q, jq) splits a range that does not); each lane copies its
// own range through xf(record, ox, oy), the four waves taking every fourth
// entry.  Every wave computes the same plan.  Returns the staged count;
// staged entries [0, core_end) come from core leaves.
template <int STAGE, class R, class NearFn, class Xf>
__device__ __forceinline__ int stage_near_batch(const R *__restrict__ src,
                                                const int32_t *__restrict__ sp, R *st, int nnear,
                                                const NearFn &nf, const Xf &xf, int &q, int &jq,
                                                int &core_end)
{
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int ql = q + lane;
    int j0 = 0, len = 0, core = 0;
    float ox = 0.f, oy = 0.f;
    uint32_t sb = 0;
    if (ql < nnear && nf(ql, sb, ox, oy, core)) {
        j0 = sp[sb] + (lane == 0 ? jq : 0);
        len = sp[sb + 1] - j0;
    }
    int incl = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
    }
    const int excl = incl - len;
    const bool in = ql < nnear && excl < STAGE;
    const int take = in ? min(len, STAGE - excl) : 0;
    const int count = __popcll(__ballot(in));  // >= 1: lane 0 is in
    const int last_take = __shfl(take, count - 1), last_len = __shfl(len, count - 1);
    const int used = __shfl(excl + take, count - 1);
    // staged entries of core leaves form a prefix of the batch
    const unsigned long long far_mask = __ballot(take > 0 && !core);
    core_end = far_mask ? __shfl(excl, __ffsll((long long)far_mask) - 1) : used;
    if (last_take < last_len) {
        jq = (count == 1 ? jq : 0) + last_take;
        q += count - 1;
    } else {
        jq = 0;
        q += count;
    }
    if (in) {
        const R *gs = src + j0;
        R *ls = st + excl;
        int e = wv;
        for (; e + 28 < take; e += 32) {
            R r[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) r[u] = gs[e + 4 * u];
#pragma unroll
            for (int u = 0; u < 8; ++u) ls[e + 4 * u] = xf(r[u], ox, oy);
        }
        for (; e < take; e += 4) ls[e] = xf(gs[e], ox, oy);
    }
    return used;
}

// float32 records shifted into the target leaf's frame and scaled by lam
template <class NearFn>
__device__ __forceinline__ int p2p_stage_batch(const float4 *__restrict__ src,
                                               const int32_t *__restrict__ sp, float4 *st,
                                               int nnear, const NearFn &nf, float lam, int &q,
                                               int &jq, int &core_end)
{
    auto xf = [lam](float4 r, float ox, float oy) {
        r.x = (r.x + ox) * lam;
        r.y = (r.y + oy) * lam;
        return r;
    };
    return stage_near_batch<P2P_STAGE>(src, sp, st, nnear, nf, xf, q, jq, core_end);
}

// Sum over the staged entries x in [a, b) with x = o (mod S) for NT
// targets (xt[u], yt[u]) at once (each staged record read once for all of
// them): per target, plain float32 partials over at most 16 pairs folded
// into the TwoSum accumulators (the same order for any NT).  S = 16 (a
// lane's share when 16 lanes serve 16 targets) or 32 (two lanes per target).
template <int KIND, bool CORE, int NT, int S = 16>
__device__ __forceinline__ void p2p_sweep(const float4 *__restrict__ st, int a, int b, int o,
                                          const float *xt, const float *yt, float c2s,
                                          float scale_f, float *ha, float *la, float *hb,
                                          float *lb)
{
    static_assert(S == 16 || S == 32, "stride");
    int x = a + ((o - a) & (S - 1));
    for (; x + 15 * S < b; x += 16 * S) {
        float aa[NT], ab[NT];
#pragma unroll
        for (int u = 0; u < NT; ++u) aa[u] = ab[u] = 0.f;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const float4 q = st[x + S * k];
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                const float dx = xt[u] - q.x, dy = yt[u] - q.y;
                const float P = pair_scaled<KIND, CORE>(dx, dy, q.z, c2s, scale_f);
                aa[u] = __builtin_fmaf(P, dx, aa[u]);
                ab[u] = __builtin_fmaf(P, dy, ab[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            two_sum_add(ha[u], la[u], aa[u]);
            two_sum_add(hb[u], lb[u], ab[u]);
        }
    }
    if (x < b) {
        float aa[NT], ab[NT];
#pragma unroll
        for (int u = 0; u < NT; ++u) aa[u] = ab[u] = 0.f;
        for (; x + 3 * S < b; x += 4 * S) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float4 q = st[x + S * k];
#pragma unroll
                for (int u = 0; u < NT; ++u) {
                    const float dx = xt[u] - q.x, dy = yt[u] - q.y;
                    const float P = pair_scaled<KIND, CORE>(dx, dy, q.z, c2s, scale_f);
                    aa[u] = __builtin_fmaf(P, dx, aa[u]);
                    ab[u] = __builtin_fmaf(P, dy, ab[u]);
                }
            }
        }
        for (; x < b; x += S) {
            const float4 q = st[x];
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                const float dx = xt[u] - q.x, dy = yt[u] - q.y;
                const float P = pair_scaled<KIND, CORE>(dx, dy, q.z, c2s, scale_f);
                aa[u] = __builtin_fmaf(P, dx, aa[u]);
                ab[u] = __builtin_fmaf(P, dy, ab[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            two_sum_add(ha[u], la[u], aa[u]);
            two_sum_add(hb[u], lb[u], ab[u]);
        }
    }
}

constexpr int P2P_CHUNKS = 8;  // target chunks of 16 per round (128 targets)

// Near-field sums of the targets [t0, t1) of one leaf, in scaled
// coordinates x' = lam x (pair_scaled).  nf(q, sb, ox, oy, core) gives near
// index q's source leaf sb, its centre shift (exact in float32) and whether
// its pairs may lie inside the smoothing core (false: no such leaf); near
// indices with core set come first.  tgt(i): target i's offset from the
// leaf centre.  Targets go in rounds of up to 128 (chunks of 16); every
// staged batch of near sources is swept by all four waves for every chunk,
// lane 16 g + t of wave w taking target t of the chunk and the staged
// entries = 4 w + g (mod 16), so the waves carry equal work whatever the
// leaf population (HALF: a last chunk of at most 8 targets has lane 16 g + t
// take target t mod 8 and the entries = 4 w + g + 16 (t / 8) (mod 32); off
// in the batched kernel, where it measured slower).  Per chunk the four groups of a wave are combined by a
// butterfly and the wave's partial is merged into its own LDS slot
// acc[(c * 4 + w) * 16 + t]; at the end of a round one thread per target
// merges the four waves' slots in wave order and calls
// out(i, hi_a, lo_a, hi_b, lo_b): the sums are lam (hi + lo).  Every thread
// of the workgroup must call this (it synchronises).
template <int KIND, bool HALF = true, class NearFn, class TgtFn, class OutFn>
__device__ __forceinline__ void p2p_leaf(const float4 *__restrict__ src,
                                         const int32_t *__restrict__ sp, float4 *st, float4 *acc,
                                         int nnear, const NearFn &nf, int32_t t0, int32_t t1,
                                         const TgtFn &tgt, const PairParams &pp, float lam,
                                         const OutFn &out)
{
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, g = lane >> 4, t = lane & 15;
    const int o = 4 * wv + g;
    const float c2s = pp.c2_f * (lam * lam);
    for (int32_t rc = t0; rc < t1; rc += 16 * P2P_CHUNKS) {
        const int nch = min(P2P_CHUNKS, (int)((t1 - rc + 15) >> 4));
        // a last chunk of 1 .. 8 targets goes alone with two lanes per
        // target (sources x = o + 16 (t >> 3) mod 32), so half its lanes
        // are not spent on absent targets
        const int rem = (int)min(t1 - rc, (int32_t)(16 * P2P_CHUNKS)) & 15;
        const bool half = HALF && rem >= 1 && rem <= 8;
        const int npair = half ? nch - 1 : nch;
        if (g == 0)
            for (int c = 0; c < nch; ++c)
                acc[(c * 4 + wv) * 16 + t] = make_float4(0.f, 0.f, 0.f, 0.f);
        int q = 0, jq = 0;
        while (q < nnear) {
            int core_end;
            const int used = p2p_stage_batch(src, sp, st, nnear, nf, lam, q, jq, core_end);
            __syncthreads();
            // chunks two at a time (each staged record read once for both)
            for (int c = 0; c < npair; c += 2) {
                float xt[2], yt[2], ha[2] = {}, la[2] = {}, hb[2] = {}, lb[2] = {};
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int32_t i = rc + 16 * (c + u) + t;
                    const float2 tp = i < t1 ? tgt(i) : make_float2(0.f, 0.f);
                    xt[u] = tp.x * lam;
                    yt[u] = tp.y * lam;
                }
                const int nu = min(2, npair - c);
                if (nu == 2) {
                    p2p_sweep<KIND, true, 2>(st, 0, core_end, o, xt, yt, c2s, pp.scale_f, ha, la,
                                             hb, lb);
                    p2p_sweep<KIND, false, 2>(st, core_end, used, o, xt, yt, c2s, pp.scale_f, ha,
                                              la, hb, lb);
                } else {
                    p2p_sweep<KIND, true, 1>(st, 0, core_end, o, xt, yt, c2s, pp.scale_f, ha, la,
                                             hb, lb);
                    p2p_sweep<KIND, false, 1>(st, core_end, used, o, xt, yt, c2s, pp.scale_f, ha,
                                              la, hb, lb);
                }
                for (int u = 0; u < nu; ++u) {
                    p2p_merge_groups(ha[u], la[u], hb[u], lb[u]);
                    if (g == 0) {
                        float4 &A = acc[((c + u) * 4 + wv) * 16 + t];
                        float4 v = A;
                        two_sum_merge(v.x, v.y, ha[u], la[u]);
                        two_sum_merge(v.z, v.w, hb[u], lb[u]);
                        A = v;
                    }
                }
            }
            if (half) {
                const int c = nch - 1, th = t & 7;
                const int32_t i = rc + 16 * c + th;
                const float2 tp = i < t1 ? tgt(i) : make_float2(0.f, 0.f);
                float xt[1] = {tp.x * lam}, yt[1] = {tp.y * lam};
                float ha[1] = {}, la[1] = {}, hb[1] = {}, lb[1] = {};
                const int oh = o + 16 * (t >> 3);
                p2p_sweep<KIND, true, 1, 32>(st, 0, core_end, oh, xt, yt, c2s, pp.scale_f, ha, la,
                                             hb, lb);
                p2p_sweep<KIND, false, 1, 32>(st, core_end, used, oh, xt, yt, c2s, pp.scale_f, ha,
                                              la, hb, lb);
                two_sum_merge(ha[0], la[0], __shfl_xor(ha[0], 8), __shfl_xor(la[0], 8));
                two_sum_merge(hb[0], lb[0], __shfl_xor(hb[0], 8), __shfl_xor(lb[0], 8));
                p2p_merge_groups(ha[0], la[0], hb[0], lb[0]);
                if (g == 0 && t < 8) {
                    float4 &A = acc[(c * 4 + wv) * 16 + t];
                    float4 v = A;
                    two_sum_merge(v.x, v.y, ha[0], la[0]);
                    two_sum_merge(v.z, v.w, hb[0], lb[0]);
                    A = v;
                }
            }
            __syncthreads();
        }
        if (tid < 16 * nch && rc + tid < t1) {
            const int c = tid >> 4, tt = tid & 15;
            float4 v = acc[(c * 4) * 16 + tt];
#pragma unroll
            for (int w = 1; w < 4; ++w) {
                const float4 u = acc[(c * 4 + w) * 16 + tt];
                two_sum_merge(v.x, v.y, u.x, u.y);
                two_sum_merge(v.z, v.w, u.z, u.w);
            }
            out(rc + tid, v.x, v.y, v.z, v.w);
        }
        __syncthreads();  // acc is reused by the next round
    }
}

// ---------------------------------------------------------------- float64
constexpr int P2P_STAGE_F64 = 768;  // staged sources (double4) per batch

// float64 pair sweep over staged entries x in [a, b) with x = o (mod S)
// for NT targets at once (each staged record read once for all of them),
// absolute coordinates, pair_f64's operation order.  (Reading the next
// entry ahead of the current pairs was measured slower: 1.62 against
// 1.55 ms at 1M blobs.)
template <int KIND, int NT, int S = 16>
__device__ __forceinline__ void p2p_sweep_f64(const double4 *__restrict__ st, int a, int b, int o,
                                              const double *xt, const double *yt,
                                              const PairParams &pp, double *ax, double *ay)
{
    for (int x = a + ((o - a) & (S - 1)); x < b; x += S) {
        const double4 q = st[x];
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            const double dx = xt[u] - q.x, dy = yt[u] - q.y;
            double P;
            if (KIND == PAIR_OSEEN || KIND == PAIR_OSEEN_NOCUT) {
                // pair_f64's Oseen form; the smoothing factor is skipped
                // (1 exactly) when no lane of the wave is inside the cutoff.
                // The 1e-200 in r2 changes no pair at a distance (it is far
                // below half an ulp of any r2 above 1e-184) and makes a
                // blob's pair with itself finite, so its P dx is 0 as
                // pair_f64's r2 > 0 test makes it, without that compare and
                // two selects per pair.
                const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, 1e-200));
                double om = 1.0;
                if (__ballot(r2 < pp.c2))
                    om = __builtin_fmin(-expm1_neg_f64(-pp.inv_ks * r2) * pp.scale, 1.0);
                P = (q.z * rcp_f64(r2)) * om;
            } else {
                P = pair_f64<KIND>(dx, dy, q.z, pp);
            }
            ax[u] = __builtin_fma(P, dx, ax[u]);
            ay[u] = __builtin_fma(P, dy, ay[u]);
        }
    }
}

// float64 near-field sums of the targets [t0, t1) of one leaf, the
// float32 p2p_leaf's layout with float64 records and plain float64
// partials: out(i, sum_a, sum_b).
template <int KIND, class NearFn, class TgtFn, class OutFn>
__device__ __forceinline__ void p2p_leaf_f64(const double4 *__restrict__ src,
                                             const int32_t *__restrict__ sp, double4 *st,
                                             double2 *acc, int nnear, const NearFn &nf,
                                             int32_t t0, int32_t t1, const TgtFn &tgt,
                                             const PairParams &pp, const OutFn &out)
{
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, g = lane >> 4, t = lane & 15;
    const int o = 4 * wv + g;
    auto xf = [](double4 r, float, float) { return r; };
    for (int32_t rc = t0; rc < t1; rc += 16 * P2P_CHUNKS) {
        const int nch = min(P2P_CHUNKS, (int)((t1 - rc + 15) >> 4));
        // a last chunk of 1 .. 8 targets: two lanes per target (as p2p_leaf)
        const int rem = (int)min(t1 - rc, (int32_t)(16 * P2P_CHUNKS)) & 15;
        const bool half = rem >= 1 && rem <= 8;
        const int npair = half ? nch - 1 : nch;
        if (g == 0)
            for (int c = 0; c < nch; ++c) acc[(c * 4 + wv) * 16 + t] = make_double2(0.0, 0.0);
        int q = 0, jq = 0;
        while (q < nnear) {
            int core_end;
            const int used = stage_near_batch<P2P_STAGE_F64>(src, sp, st, nnear, nf, xf, q, jq,
                                                             core_end);
            __syncthreads();
            // chunks two at a time (each staged record read once for both)
            for (int c = 0; c < npair; c += 2) {
                double xt[2], yt[2], ax[2] = {}, ay[2] = {};
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int32_t i = rc + 16 * (c + u) + t;
                    const double2 tp = i < t1 ? tgt(i) : make_double2(0.0, 0.0);
                    xt[u] = tp.x;
                    yt[u] = tp.y;
                }
                const int nu = min(2, npair - c);
                if (nu == 2)
                    p2p_sweep_f64<KIND, 2>(st, 0, used, o, xt, yt, pp, ax, ay);
                else
                    p2p_sweep_f64<KIND, 1>(st, 0, used, o, xt, yt, pp, ax, ay);
                for (int u = 0; u < nu; ++u) {
                    ax[u] += __shfl_xor(ax[u], 16);
                    ay[u] += __shfl_xor(ay[u], 16);
                    ax[u] += __shfl_xor(ax[u], 32);
                    ay[u] += __shfl_xor(ay[u], 32);
                    if (g == 0) {
                        double2 &A = acc[((c + u) * 4 + wv) * 16 + t];
                        A = make_double2(A.x + ax[u], A.y + ay[u]);
                    }
                }
            }
            if (half) {
                const int c = nch - 1, th = t & 7;
                const int32_t i = rc + 16 * c + th;
                const double2 tp = i < t1 ? tgt(i) : make_double2(0.0, 0.0);
                double xt[1] = {tp.x}, yt[1] = {tp.y}, ax[1] = {}, ay[1] = {};
                p2p_sweep_f64<KIND, 1, 32>(st, 0, used, o + 16 * (t >> 3), xt, yt, pp, ax, ay);
                ax[0] += __shfl_xor(ax[0], 8);
                ay[0] += __shfl_xor(ay[0], 8);
                ax[0] += __shfl_xor(ax[0], 16);
                ay[0] += __shfl_xor(ay[0], 16);
                ax[0] += __shfl_xor(ax[0], 32);
                ay[0] += __shfl_xor(ay[0], 32);
                if (g == 0 && t < 8) {
                    double2 &A = acc[(c * 4 + wv) * 16 + t];
                    A = make_double2(A.x + ax[0], A.y + ay[0]);
                }
            }
            __syncthreads();
        }
        if (tid < 16 * nch && rc + tid < t1) {
            const int c = tid >> 4, tt = tid & 15;
            double2 v = acc[(c * 4) * 16 + tt];
#pragma unroll
            for (int w = 1; w < 4; ++w) {
                const double2 u = acc[(c * 4 + w) * 16 + tt];
                v.x += u.x;
                v.y += u.y;
            }
            out(rc + tid, v.x, v.y);
        }
        __syncthreads();  // acc is reused by the next round
    }
}

}  // namespace phf
