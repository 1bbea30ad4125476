// phf_p2p_dev.hpp -- the float32 near-field sum over LDS-staged sources,
// shared by the single-system and the batched FMM.
//
// A workgroup of 4 waves serves one target leaf.  The sources of the
// leaf's near leaves (ranges padded to multiples of 4 with zero-weight
// entries) are staged in LDS in batches: lane l of every wave looks at near
// index q + l, a wave scan of the range lengths packs as many as fit (a
// cursor splits a range that does not), and each lane copies its own range
// with 8 loads in flight, the four waves taking every fourth entry.  Lane
// 16 g + t then sums staged sources g, g+4, g+8, ... for target t.
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

template <int KIND, bool CORE>
__device__ __forceinline__ void p2p_seg_f32(const float4 *__restrict__ sq, int j0, int j1,
                                            float xt, float yt, float c2s, float scale_f,
                                            float &ha, float &la, float &hb, float &lb)
{
    int j = j0;
    for (; j + 64 <= j1; j += 64) {
        float aa = 0.f, ab = 0.f;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const float4 q = sq[j + 4 * k];
            const float dx = xt - q.x, dy = yt - q.y;
            const float P = pair_scaled<KIND, CORE>(dx, dy, q.z, c2s, scale_f);
            aa = __builtin_fmaf(P, dx, aa);
            ab = __builtin_fmaf(P, dy, ab);
        }
        two_sum_add(ha, la, aa);
        two_sum_add(hb, lb, ab);
    }
    if (j < j1) {
        float aa = 0.f, ab = 0.f;
        for (; j + 16 <= j1; j += 16) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float4 q = sq[j + 4 * k];
                const float dx = xt - q.x, dy = yt - q.y;
                const float P = pair_scaled<KIND, CORE>(dx, dy, q.z, c2s, scale_f);
                aa = __builtin_fmaf(P, dx, aa);
                ab = __builtin_fmaf(P, dy, ab);
            }
        }
        for (; j < j1; j += 4) {
            const float4 q = sq[j];
            const float dx = xt - q.x, dy = yt - q.y;
            const float P = pair_scaled<KIND, CORE>(dx, dy, q.z, c2s, scale_f);
            aa = __builtin_fmaf(P, dx, aa);
            ab = __builtin_fmaf(P, dy, ab);
        }
        two_sum_add(ha, la, aa);
        two_sum_add(hb, lb, ab);
    }
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

// Accumulate the near-field sum of target tp (offset from its leaf centre),
// in scaled coordinates: the true sums are lam_out times (ha + la), (hb + lb)
// over every near leaf nf names: nf(q, sb, ox, oy, core) gives near index
// q's source leaf sb, its centre shift (exact in float32) and whether its
// pairs may lie inside the smoothing core; false: no such leaf.  Near
// indices with core set must come first.  Sources are staged shifted into
// the target leaf's frame, so each batch is two contiguous ranges (core,
// then beyond the core) summed without per-leaf setup.  Every thread of
// the workgroup must call this (it synchronises); wave_on: the wave has
// targets.
template <int KIND, class NearFn>
__device__ __forceinline__ void p2p_staged(const float4 *__restrict__ src,
                                           const int32_t *__restrict__ sp, float4 *st,
                                           int nnear, const NearFn &nf,
                                           bool wave_on, float2 tp, const PairParams &pp,
                                           float &ha, float &la, float &hb, float &lb,
                                           float &lam_out)
{
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63, g = lane >> 4;
    // scaled coordinates (pair_scaled): sources are scaled while staged
    const float lam = __builtin_sqrtf(-pp.nlog2e_ks_f);
    const float c2s = pp.c2_f * (lam * lam);
    const float txs = tp.x * lam, tys = tp.y * lam;
    lam_out = lam;
    int q = 0, jq = 0;  // cursor: near index, position within its range
    while (q < nnear) {
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
        const bool in = ql < nnear && excl < P2P_STAGE;
        const int take = in ? min(len, P2P_STAGE - excl) : 0;  // multiple of 4
        const int count = __popcll(__ballot(in));
        const int last_take = __shfl(take, count - 1), last_len = __shfl(len, count - 1);
        const int used = __shfl(excl + take, count - 1);
        // staged entries of core leaves form a prefix of the batch
        const unsigned long long far_mask = __ballot(take > 0 && !core);
        const int core_end = far_mask ? __shfl(excl, __ffsll((long long)far_mask) - 1) : used;
        if (last_take < last_len) {
            jq = (count == 1 ? jq : 0) + last_take;
            q += count - 1;
        } else {
            jq = 0;
            q += count;
        }
        if (in) {
            const float4 *gs = src + j0;
            float4 *ls = st + excl;
            int e = wv;
            for (; e + 28 < take; e += 32) {
                float4 r[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) r[u] = gs[e + 4 * u];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    r[u].x = (r[u].x + ox) * lam;
                    r[u].y = (r[u].y + oy) * lam;
                    ls[e + 4 * u] = r[u];
                }
            }
            for (; e < take; e += 4) {
                float4 r = gs[e];
                r.x = (r.x + ox) * lam;
                r.y = (r.y + oy) * lam;
                ls[e] = r;
            }
        }
        __syncthreads();
        if (wave_on) {
            const float4 *sq = st + g;
            p2p_seg_f32<KIND, true>(sq, 0, core_end, txs, tys, c2s, pp.scale_f, ha, la, hb, lb);
            p2p_seg_f32<KIND, false>(sq, core_end, used, txs, tys, c2s, pp.scale_f, ha, la, hb,
                                     lb);
        }
        __syncthreads();
    }
}

}  // namespace phf
