// phf_panel.hip -- vortex-panel influence matrix, panel-induced velocity
// and the dense panel solve on gfx950.
//
// Replaces
//   assemble_influenceMatrix / inducedVelocity
//     src/cpp/panels/cython/vortexPanel_Cython/_vortexPanel_Cython_cpu.pyx:81-233, 243-391
//   Panels.solve, src/python/panels/__panelsClass.py:715-822:
//     'fixed'      scipy lu_factor/lu_solve (LAPACK getrf/getrs)   :961, :782
//     'prescribed' numpy.linalg.lstsq([A; L], [RHS; -gTotal])      :800-822
//
// Solve = LU factorisation in float32 with partial pivoting (blocked,
// right-looking; the trailing update runs on MFMA, v_mfma_f32_16x16x4f32)
// + iterative refinement with float64 residuals until the correction stalls
// at float64 resolution.  Over-determined systems (the 'prescribed' problem)
// are solved through the normal equations B^T B x = B^T c, refined with the
// float64 semi-normal residual B^T (c - B x).
#include "phf_panel.hpp"
#include "phf_tree.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <type_traits>
#include <vector>

namespace phf {

// ---------------------------------------------------- influence matrix ----
// A[i*M + j]: tangential velocity at collocation point i induced by unit
// strength on panel j (pyx:208-233, same operation order).
__global__ void panel_influence_kernel(int64_t M, const double *__restrict__ xCP,
                                       const double *__restrict__ yCP,
                                       const double *__restrict__ xs, const double *__restrict__ ys,
                                       const double *__restrict__ xe, const double *__restrict__ ye,
                                       const double *__restrict__ cosA,
                                       const double *__restrict__ sinA, double *__restrict__ A)
{
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * M) return;
    const int64_t i = t / M, j = t - (t / M) * M;
    const double inv_2pi = 1.0 / (2.0 * M_PI);
    const double inv_4pi = inv_2pi / 2.0;
    const double Lx = xe[j] - xs[j];
    const double Ly = ye[j] - ys[j];
    const double x2 = cosA[j] * Lx + sinA[j] * Ly;
    const double xP = cosA[j] * (xCP[i] - xs[j]) + sinA[j] * (yCP[i] - ys[j]);
    const double yP = -sinA[j] * (xCP[i] - xs[j]) + cosA[j] * (yCP[i] - ys[j]);
    const double vxP = inv_2pi * (atan2(yP, (xP - x2)) - atan2(yP, xP));
    const double vyP = -inv_4pi * log((xP * xP + yP * yP) / ((xP - x2) * (xP - x2) + yP * yP));
    const double vx = cosA[j] * vxP - sinA[j] * vyP;
    const double vy = sinA[j] * vxP + cosA[j] * vyP;
    A[t] = vx * cosA[i] + vy * sinA[i];
}

// Velocity induced at evaluation points by panels of strength gamma
// (pyx:357-391).  Per panel, in its frame (xP along the panel from its
// start, yP normal; panel length x2):
//   u = gamma/(2 pi) (atan2(yP, xP - x2) - atan2(yP, xP))
//   v = -gamma/(4 pi) log((xP^2 + yP^2) / ((xP - x2)^2 + yP^2))
// rotated back to the global frame.  Evaluated here as
//   atan2(yP x2, xP (xP - x2) + yP^2)        (the angle between the two
//                                              endpoint vectors; both
//                                              atan2 share the sign of yP,
//                                              so the difference lies in
//                                              [-pi, pi] and the forms agree)
//   log1p(x2 (2 xP - x2) / ((xP - x2)^2 + yP^2))   (num - den formed exactly)
// which avoids the cancellation of the original forms far from the panel
// (one atan2 and one log1p per pair instead of two atan2, a log and a
// division).  Panel constants are formed once and staged in LDS.
struct PanelK {
    double xs, ys, c, s;  // start, cos, sin
    double x2, g2, g4, pad;  // length along the panel, gamma/(2 pi), gamma/(4 pi)
};

__global__ void panel_consts(int64_t M, const double *__restrict__ gamma,
                             const double *__restrict__ xs, const double *__restrict__ ys,
                             const double *__restrict__ xe, const double *__restrict__ ye,
                             const double *__restrict__ cosA, const double *__restrict__ sinA,
                             PanelK *__restrict__ out)
{
#pragma clang fp contract(off)
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const double inv_2pi = 1.0 / (2.0 * M_PI);
    const double inv_4pi = inv_2pi / 2.0;
    const double Lx = xe[j] - xs[j];
    const double Ly = ye[j] - ys[j];
    PanelK k;
    k.xs = xs[j];
    k.ys = ys[j];
    k.c = cosA[j];
    k.s = sinA[j];
    k.x2 = cosA[j] * Lx + sinA[j] * Ly;
    k.g2 = gamma[j] * inv_2pi;
    k.g4 = gamma[j] * inv_4pi;
    k.pad = 0.0;
    out[j] = k;
}

constexpr int PI_T = 256;      // evaluation points per block
constexpr int PI_CHUNK = 256;  // panels staged per round

// F32 = false: float64 throughout.  F32 = true (mixed-precision steps): the
// panel-frame coordinates and the arguments are formed in float64, the
// atan2 / log1p in float32, sums in float64 (about 1e-7 relative per pair).
template <bool F32>
__global__ __launch_bounds__(PI_T) void panel_induced_kernel(int64_t M, const PanelK *__restrict__ pk,
                                                              int64_t ne,
                                                              const double *__restrict__ xE,
                                                              const double *__restrict__ yE,
                                                              double *__restrict__ vx,
                                                              double *__restrict__ vy)
{
    __shared__ PanelK sh[PI_CHUNK];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * PI_T + tid;
    const double x = i < ne ? xE[i] : 0.0, y = i < ne ? yE[i] : 0.0;
    double sx = 0.0, sy = 0.0;
    for (int64_t j0 = 0; j0 < M; j0 += PI_CHUNK) {
        const int cnt = (int)min((int64_t)PI_CHUNK, M - j0);
        __syncthreads();
        if (tid < cnt) sh[tid] = pk[j0 + tid];
        __syncthreads();
        for (int k = 0; k < cnt; ++k) {
            const PanelK p = sh[k];
            const double dx = x - p.xs, dy = y - p.ys;
            const double xP = p.c * dx + p.s * dy;
            const double yP = -p.s * dx + p.c * dy;
            const double xm = xP - p.x2;
            const double cross = yP * p.x2, dot = xP * xm + yP * yP;
            const double den = xm * xm + yP * yP;
            const double num = p.x2 * (2.0 * xP - p.x2);
            double th, lg;
            if (F32) {
                th = (double)atan2f((float)cross, (float)dot);
                lg = (double)log1pf((float)num / (float)den);
            } else {
                th = atan2(cross, dot);
                lg = log1p(num / den);
            }
            const double vxP = p.g2 * th;
            const double vyP = -p.g4 * lg;
            sx += p.c * vxP - p.s * vyP;
            sy += p.s * vxP + p.c * vyP;
        }
    }
    if (i < ne) {
        vx[i] = sx;
        vy[i] = sy;
    }
}

// ------------------------------------------------------- matrix setup ----
// F64 row-major B (rows x n) -> F64 column-major copy (for residuals) and,
// for square systems, the F32 column-major working matrix.
__global__ void to_colmajor(const double *__restrict__ B, int64_t rows, int64_t n,
                            double *__restrict__ Bc, float *__restrict__ A32)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * n) return;
    const int64_t j = t / rows, i = t - (t / rows) * rows;  // column-major index t = j*rows + i
    const double v = B[i * n + j];
    Bc[t] = v;
    if (A32) A32[t] = (float)v;
}

// N = B^T B (n x n, column-major, F64 and F32 copies); Bc column-major rows x n.
// One workgroup per 32 x 32 tile of N: the 32 columns of B on each side are
// staged in LDS 32 rows at a time and each thread adds 2 x 2 outputs, every
// output summed over the rows in order (one FMA chain).
constexpr int NM_T = 32;
__global__ __launch_bounds__(256) void normal_matrix(const double *__restrict__ Bc, int64_t rows,
                                                     int64_t n, double *__restrict__ N64,
                                                     float *__restrict__ N32)
{
    __shared__ double Ti[NM_T][NM_T + 1];  // [column of the i side][row]
    __shared__ double Tj[NM_T][NM_T + 1];
    const int64_t i0 = (int64_t)blockIdx.x * NM_T, j0 = (int64_t)blockIdx.y * NM_T;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int64_t r0 = 0; r0 < rows; r0 += NM_T) {
        for (int e = threadIdx.x; e < NM_T * NM_T; e += 256) {
            const int c = e / NM_T, k = e % NM_T;
            const int64_t r = r0 + k;
            Ti[c][k] = (i0 + c < n && r < rows) ? Bc[(i0 + c) * rows + r] : 0.0;
            Tj[c][k] = (j0 + c < n && r < rows) ? Bc[(j0 + c) * rows + r] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < NM_T; ++k) {
            const double a0 = Ti[tx][k], a1 = Ti[tx + 16][k];
            const double b0 = Tj[ty][k], b1 = Tj[ty + 16][k];
            acc[0][0] = __builtin_fma(a0, b0, acc[0][0]);
            acc[1][0] = __builtin_fma(a1, b0, acc[1][0]);
            acc[0][1] = __builtin_fma(a0, b1, acc[0][1]);
            acc[1][1] = __builtin_fma(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int64_t i = i0 + tx + 16 * a, j = j0 + ty + 16 * b;
            if (i < n && j < n) {
                N64[j * n + i] = acc[a][b];
                N32[j * n + i] = (float)acc[a][b];
            }
        }
}

// --------------------------------------------------------------- LU -------
constexpr int NB = 32;  // panel width

// Unblocked LU with partial pivoting of the panel A[k0:n, k0:k0+NB] (column
// major, lda = n) in LDS.  One workgroup of 256 threads; rows <= 1024.
__global__ __launch_bounds__(256) void lu_panel(float *__restrict__ A, int n, int k0,
                                                int32_t *__restrict__ ipiv,
                                                int32_t *__restrict__ info, int use_lds)
{
    extern __shared__ float lds_panel[];  // [NB][rows] column-major, rows = n - k0
    __shared__ float red_v[256];
    __shared__ int red_i[256];
    const int rows = n - k0;
    const int w = min(NB, n - k0);
    const int tid = threadIdx.x;
    // the panel is worked on in LDS when it fits, else in place (L2-resident)
    float *P = use_lds ? lds_panel : A + (int64_t)k0 * n + k0;
    const int64_t ld = use_lds ? rows : n;
    if (use_lds)
        for (int c = 0; c < w; ++c)
            for (int r = tid; r < rows; r += 256) P[c * ld + r] = A[(int64_t)(k0 + c) * n + k0 + r];
    __syncthreads();
    for (int c = 0; c < w; ++c) {
        // pivot: first row of max |P[r][c]|, r >= c
        float best = -1.0f;
        int bi = c;
        for (int r = c + tid; r < rows; r += 256) {
            const float v = fabsf(P[c * ld + r]);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        red_v[tid] = best;
        red_i[tid] = bi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s) {
                const float a = red_v[tid], b = red_v[tid + s];
                const int ia = red_i[tid], ib = red_i[tid + s];
                if (b > a || (b == a && ib < ia)) {
                    red_v[tid] = b;
                    red_i[tid] = ib;
                }
            }
            __syncthreads();
        }
        const int p = red_i[0];
        const float pv = red_v[0];
        __syncthreads();
        if (tid == 0) {
            ipiv[k0 + c] = k0 + p;
            if (pv == 0.0f && info[0] == 0) info[0] = k0 + c + 1;
        }
        if (p != c)
            for (int cc = tid; cc < w; cc += 256) {
                const float t = P[cc * ld + c];
                P[cc * ld + c] = P[cc * ld + p];
                P[cc * ld + p] = t;
            }
        __syncthreads();
        const float d = P[c * ld + c];
        if (d != 0.0f) {
            const float inv = 1.0f / d;
            for (int r = c + 1 + tid; r < rows; r += 256) P[c * ld + r] *= inv;
        }
        __syncthreads();
        for (int cc = c + 1; cc < w; ++cc) {
            const float u = P[cc * ld + c];
            for (int r = c + 1 + tid; r < rows; r += 256)
                P[cc * ld + r] = __builtin_fmaf(-P[c * ld + r], u, P[cc * ld + r]);
        }
        __syncthreads();
    }
    if (use_lds)
        for (int c = 0; c < w; ++c)
            for (int r = tid; r < rows; r += 256) A[(int64_t)(k0 + c) * n + k0 + r] = P[c * ld + r];
}

// Row swaps of panel k0 applied to every column outside the panel, then
// U12 = L11^-1 A12 for the columns right of it.  One thread per column.
__global__ void lu_swap_trsm(float *__restrict__ A, int n, int k0,
                             const int32_t *__restrict__ ipiv)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int w = min(NB, n - k0);
    if (j >= n || (j >= k0 && j < k0 + w)) return;
    float *col = A + (int64_t)j * n;
    for (int c = 0; c < w; ++c) {
        const int p = ipiv[k0 + c];
        if (p != k0 + c) {
            const float t = col[k0 + c];
            col[k0 + c] = col[p];
            col[p] = t;
        }
    }
    if (j < k0 + w) return;
    float u[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) u[c] = c < w ? col[k0 + c] : 0.0f;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
#pragma unroll
        for (int r = c + 1; r < NB; ++r) {
            const float l = (r < w) ? A[(int64_t)(k0 + c) * n + k0 + r] : 0.0f;
            u[r] = __builtin_fmaf(-l, u[c], u[r]);
        }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c)
        if (c < w) col[k0 + c] = u[c];
}

// Trailing update A22 -= L21 * U12 with MFMA: one wave per 16x16 tile of
// A22, K = NB in steps of 4 (v_mfma_f32_16x16x4f32; A operand lane l holds
// A[l%16][l/16], B operand B[l/16][l%16], result D[4*(l/16)+v][l%16]).
typedef float floatx4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(64) void lu_update_mfma(float *__restrict__ A, int n, int k0)
{
    const int r0 = k0 + NB, c0 = k0 + NB;
    const int m = n - r0;
    const int tiles_r = (m + 15) / 16;
    const int tr = blockIdx.x % tiles_r, tc = blockIdx.x / tiles_r;
    const int lane = threadIdx.x;
    const int li = lane & 15, lk = lane >> 4;
    const int row = r0 + tr * 16, col = c0 + tc * 16;
    floatx4 acc;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = row + 4 * lk + v, j = col + li;
        acc[v] = (i < n && j < n) ? A[(int64_t)j * n + i] : 0.0f;
    }
#pragma unroll
    for (int kk = 0; kk < NB; kk += 4) {
        const int i = row + li, j = col + li, k = k0 + kk + lk;
        const float a = (i < n) ? A[(int64_t)k * n + i] : 0.0f;   // L21[i][k]
        const float b = (j < n) ? A[(int64_t)j * n + k] : 0.0f;   // U12[k][j]
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(-a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = row + 4 * lk + v, j = col + li;
        if (i < n && j < n) A[(int64_t)j * n + i] = acc[v];
    }
}

// -------------------------------------------- blocked LU, n <= 1024 ------
// Right-looking LU with partial pivoting in panels of NB columns, one
// launch per panel step, the next panel factored inside the same launch:
//   lu_first     one workgroup factors panel 0;
//   lu_step(k)   one workgroup per NB-column block j of the matrix:
//                  j <  k   the row swaps of panel k on its columns (L part)
//                  j == k   nothing (factored by the launch before)
//                  j >  k   the swaps, U12 = L11^-1 A12, A22 -= L21 U12 on
//                           the matrix cores, and for j == k + 1 the
//                           factorisation of panel k + 1.
// A panel is factored with its rows in registers (row k1 + t in thread t):
// per column a DPP maximum in each wave, its row by ballot, and one LDS
// barrier after which a single 64-bit LDS maximum names the pivot row and
// the wave whose slot holds its entries (double-buffered by column
// parity, so the next column needs no second barrier).  Each row
// carries the index of the row it came from; the rows that moved give the
// panel's composite permutation (at most 2 NB rows), which the other blocks
// apply with one load and one store per element.
// ipiv follows LAPACK getrf: row k was exchanged with row ipiv[k] at step k.
constexpr int LU_TMAX = 1024;  // n <= LU_TMAX: one row per thread

struct LuSwaps {
    int32_t cnt, pad;
    int2 mv[2 * NB];  // (row, row its content comes from)
};

// Workgroup barrier that orders LDS accesses only: outstanding global loads
// stay in flight across it (a __syncthreads also waits for those).
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <int T>
struct LuShared {
    // per column parity: each wave's candidate row's entries of columns
    // c .. c + NB - 1
    float4 cand[2][T / 64][NB / 4];
    int piv[NB];
    // per column c mod 3: the largest (|entry|, -row, wave) over the waves
    unsigned long long key[3];
    int first_zero, cnt;
    float Lt[NB][NB + 1];  // L11, [column][row]
    float Ut[NB][NB + 1];  // U12, [column][row]
};

// Largest value of the wave (v >= -1), in every lane: DPP shifts within
// rows of 16 lanes, then row broadcasts to lane 63, then a read of lane 63.
__device__ __forceinline__ float wave_max(float v)
{
    // dpp controls: row_shr:n = 0x110 + n, row_bcast:15 = 0x142, row_bcast:31 = 0x143
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                            0x111, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                            0x112, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                            0x114, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                            0x118, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                            0x142, 0xa, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v),
                                                            0x143, 0xc, 0xf, false)));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Panel columns k1 .. k1 + w of rows k1 .. n - 1, one row per thread in
// registers, held rotated: a[j] is the entry of column c + j at step c, so
// every register index is fixed (the rotation is NB register moves per
// column, cheaper than the instruction-cache misses of a fully unrolled
// loop).  Rows are exchanged by label only: each
// thread keeps its content and carries the row index it currently
// occupies.  Column c of a content is final after step c and is stored at
// the content's own (starting) row; once the panel is done, the contents
// that ended in another row are moved there.  One LDS barrier per column:
// each wave publishes its largest entry with that entry's row before it
// (the column stores stay in flight).
template <int T>
__device__ __forceinline__ void lu_factor_panel(float *__restrict__ A, int n, int k1,
                                                int32_t *__restrict__ ipiv,
                                                int32_t *__restrict__ info,
                                                LuSwaps *__restrict__ sw, LuShared<T> &sh)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int w = min(NB, n - k1);
    const int r0 = k1 + tid;  // this thread's content: its starting row
    const bool live = r0 < n;
    int r = r0;               // the row it occupies now
    float a[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) a[j] = (live && j < w) ? A[(int64_t)(k1 + j) * n + r0] : 0.0f;
    if (tid == 0) {
        sh.first_zero = -1;
        sh.cnt = 0;
        sh.key[0] = sh.key[1] = sh.key[2] = 0ull;
    }
    __syncthreads();
    // columns c >= w (last panel) run on zero entries and change nothing;
    // the loop stays rolled (its body stays in the instruction cache)
#pragma unroll 1
    for (int c = 0; c < NB; ++c) {
        const int rc = k1 + c;
        const int buf = c & 1;
        // the wave's largest |a[0]| over rows >= rc and the first row with it
        const bool cand = live && r >= rc;
        const float mine = cand ? fabsf(a[0]) : -1.0f;
        const float v = wave_max(mine);
        // the smallest occupied row among the lanes holding the maximum: a
        // ballot names it when one lane holds the maximum (ties need the
        // cross-lane minimum)
        const uint64_t hold = __ballot(cand && mine == v);
        int best_r = 0x7fffffff;
        if (__popcll(hold) == 1) {
            best_r = __builtin_amdgcn_readlane(r, (int)__builtin_ctzll(hold));
        } else if (hold != 0ull) {
            best_r = cand && mine == v ? r : 0x7fffffff;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) best_r = min(best_r, __shfl_xor(best_r, o));
        }
        float pv;
        int p, pw;
        if (T == 64) {
            if (r == best_r) {
#pragma unroll
                for (int j = 0; j < NB / 4; ++j)
                    sh.cand[buf][0][j] =
                        make_float4(a[4 * j], a[4 * j + 1], a[4 * j + 2], a[4 * j + 3]);
            }
            lds_barrier();
            pv = best_r == 0x7fffffff ? -1.0f : v;
            p = best_r;
            pw = 0;
        } else {
            // the waves' candidates meet in one 64-bit LDS maximum: |value|
            // bits, then the row inverted (a tie goes to the smaller row),
            // then the wave; the candidate's row entries go to its wave's slot
            if (r == best_r) {
                const unsigned long long key =
                    ((unsigned long long)__float_as_uint(v) << 32) |
                    ((unsigned long long)(0xFFFFFFu - (uint32_t)r) << 8) | (unsigned long long)wv;
                atomicMax(&sh.key[c % 3], key);
#pragma unroll
                for (int j = 0; j < NB / 4; ++j)
                    sh.cand[buf][wv][j] =
                        make_float4(a[4 * j], a[4 * j + 1], a[4 * j + 2], a[4 * j + 3]);
            }
            lds_barrier();
            const unsigned long long key = sh.key[c % 3];
            // column c + 2's slot: last read in column c - 1, before this barrier
            if (tid == 0) sh.key[(c + 2) % 3] = 0ull;
            if (key == 0ull) {
                pv = -1.0f;
                p = 0x7fffffff;
                pw = 0;
            } else {
                pv = __uint_as_float((uint32_t)(key >> 32));
                p = (int)(0xFFFFFFu - (uint32_t)((key >> 8) & 0xFFFFFFu));
                pw = (int)(key & 0xFFu);
            }
        }
        if (tid == 0) {
            sh.piv[c] = p;
            if (pv == 0.0f && sh.first_zero < 0 && c < w) sh.first_zero = rc;
        }
        // the exchange of rows rc and p: labels only
        if (r == p)
            r = rc;
        else if (r == rc)
            r = p;
        const float4 *pr = sh.cand[buf][pw];
        float prow[NB];
#pragma unroll
        for (int j = 0; j < NB / 4; ++j) {
            const float4 q4 = pr[j];
            prow[4 * j] = q4.x;
            prow[4 * j + 1] = q4.y;
            prow[4 * j + 2] = q4.z;
            prow[4 * j + 3] = q4.w;
        }
        const float d = prow[0];
        if (d != 0.0f && r > rc && live) {
            const float l = a[0] * __builtin_amdgcn_rcpf(d);
            a[0] = l;
#pragma unroll
            for (int j = 1; j < NB; ++j) a[j] = __builtin_fmaf(-l, prow[j], a[j]);
        }
        if (live && c < w) A[(int64_t)(k1 + c) * n + r0] = a[0];
#pragma unroll
        for (int j = 0; j + 1 < NB; ++j) a[j] = a[j + 1];
        a[NB - 1] = 0.0f;
    }
    // contents that ended in another row move there (all stores done first)
    const bool moved = live && r != r0;
    __syncthreads();
    float t[NB];
    if (moved) {
#pragma unroll
        for (int c = 0; c < NB; ++c) t[c] = c < w ? A[(int64_t)(k1 + c) * n + r0] : 0.0f;
        const int q = atomicAdd(&sh.cnt, 1);
        sw->mv[q] = make_int2(r, r0);
    }
    __syncthreads();
    if (moved) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
            if (c < w) A[(int64_t)(k1 + c) * n + r] = t[c];
    }
    if (tid < w) ipiv[k1 + tid] = sh.piv[tid];
    if (tid == 0) {
        sw->cnt = sh.cnt;
        if (sh.first_zero >= 0 && info[0] == 0) info[0] = sh.first_zero + 1;
    }
}

// Rows r0 .. n - 1 of column block c0 .. c0 + wj: A -= P Q, P the w
// columns k0 .. k0 + w of A on those rows, Q[kk][c] = Ut[c][kk] (w x wj) in
// LDS.  16 x 16 tiles on v_mfma_f32_16x16x4f32 (A operand lane l holds
// A[l%16][l/16], B operand B[l/16][l%16], result D[4 (l/16) + v][l%16]).
// A wave takes whole 16-row strips (at most LU_STRIPS; n - r0 <= T: T
// covers the rows) and both 16-column tiles of each, so a strip's P
// operand is loaded once for the two tiles.  A tile's result rows
// 4 (l/16) .. + 3 of column l%16 are consecutive in the column-major A, so
// the block is read and written 16 bytes per lane where A's columns are
// 16-byte aligned.  Every strip's operands are loaded before the first
// product, so the loads of all strips are in flight together.
// rows (nullable): the results go to LDS instead, row-major with stride
// NB + 1 from row r0 (for a panel factor that follows in the same launch).
// mid() runs after the operand loads are issued and before the products
// (it may fill Ut and wait at a barrier while those loads are in flight).
struct NoMid {
    __device__ void operator()() const {}
};

template <int T, class Mid = NoMid>
__device__ __forceinline__ void lu_block_update(float *__restrict__ A, int n, int r0, int c0,
                                                int wj, int k0, int w,
                                                const float (*__restrict__ Ut)[NB + 1],
                                                float *__restrict__ rows = nullptr,
                                                Mid mid = Mid())
{
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int m = n - r0;
    if (m <= 0) {
        mid();
        return;
    }
    static_assert(NB == 32, "two 16-column tiles per block");
    constexpr int LU_STRIPS = 4, WAVES = T / 64;
    const int li = lane & 15, lk = lane >> 4;
    const int strips = (m + 15) / 16;
    const bool vec = (n & 3) == 0 && (r0 & 3) == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
    float av[LU_STRIPS][NB / 4];
    floatx4 acc[LU_STRIPS][2];
#pragma unroll
    for (int q = 0; q < LU_STRIPS; ++q) {
        const int row = r0 + 16 * (wv + q * WAVES);
        const bool on = wv + q * WAVES < strips;
        const int i0 = row + 4 * lk;
#pragma unroll
        for (int tc = 0; tc < 2; ++tc) {
            const int jj = 16 * tc + li;
            const float *src = A + (int64_t)(c0 + jj) * n + i0;
            const bool col = on && jj < wj;
            if (col && vec && i0 + 3 < n) {
                const float4 v4 = *reinterpret_cast<const float4 *>(src);
                acc[q][tc][0] = v4.x;
                acc[q][tc][1] = v4.y;
                acc[q][tc][2] = v4.z;
                acc[q][tc][3] = v4.w;
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[q][tc][v] = (col && i0 + v < n) ? src[v] : 0.0f;
            }
        }
#pragma unroll
        for (int kk = 0; kk < NB / 4; ++kk) {
            const int i = row + li, kc = 4 * kk + lk;
            av[q][kk] = (on && i < n && kc < w) ? A[(int64_t)(k0 + kc) * n + i] : 0.0f;
        }
    }
    mid();
#pragma unroll
    for (int q = 0; q < LU_STRIPS; ++q) {
        if (wv + q * WAVES >= strips) break;  // uniform across the wave
        const int row = r0 + 16 * (wv + q * WAVES);
        const int i0 = row + 4 * lk;
#pragma unroll
        for (int tc = 0; tc < 2; ++tc) {
            if (16 * tc >= wj) break;  // uniform
            const int jj = 16 * tc + li;
#pragma unroll
            for (int kk = 0; kk < NB / 4; ++kk) {
                const int kc = 4 * kk + lk;
                const float b = (kc < w && jj < wj) ? Ut[jj][kc] : 0.0f;
                acc[q][tc] = __builtin_amdgcn_mfma_f32_16x16x4f32(-av[q][kk], b, acc[q][tc], 0, 0,
                                                                  0);
            }
            if (jj >= wj) continue;
            if (rows) {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (i0 + v < n) rows[(i0 + v - r0) * (NB + 1) + jj] = acc[q][tc][v];
            } else {
                float *dst = A + (int64_t)(c0 + jj) * n + i0;
                if (vec && i0 + 3 < n) {
                    *reinterpret_cast<float4 *>(dst) =
                        make_float4(acc[q][tc][0], acc[q][tc][1], acc[q][tc][2], acc[q][tc][3]);
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (i0 + v < n) dst[v] = acc[q][tc][v];
                }
            }
        }
    }
}

template <int T>
__global__ __launch_bounds__(T) void lu_first(float *__restrict__ A, int n,
                                              int32_t *__restrict__ ipiv,
                                              int32_t *__restrict__ info,
                                              LuSwaps *__restrict__ sw)
{
    __shared__ LuShared<T> sh;
    if (threadIdx.x == 0) info[0] = 0;
    __syncthreads();
    lu_factor_panel<T>(A, n, 0, ipiv, info, sw, sh);
}

template <int T>
__global__ __launch_bounds__(T) void lu_step(float *__restrict__ A, int n, int k,
                                             int32_t *__restrict__ ipiv,
                                             int32_t *__restrict__ info,
                                             LuSwaps *__restrict__ sw)
{
    __shared__ LuShared<T> sh;
    const int j = blockIdx.x;
    if (j == k) return;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int k0 = k * NB, w = min(NB, n - k0);
    const int c0 = j * NB, wj = min(NB, n - c0);
    // 1. the composite row permutation of panel k on columns c0 .. c0 + wj
    {
        const LuSwaps &S = sw[k];
        const int cnt = S.cnt;
        constexpr int PER = (2 * NB * NB + T - 1) / T;
        float v[PER];
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const int idx = tid + T * it, q = idx / NB, cc = idx % NB;
            v[it] = (q < cnt && cc < wj) ? A[(int64_t)(c0 + cc) * n + S.mv[q].y] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const int idx = tid + T * it, q = idx / NB, cc = idx % NB;
            if (q < cnt && cc < wj) A[(int64_t)(c0 + cc) * n + S.mv[q].x] = v[it];
        }
    }
    if (j < k) return;
    __syncthreads();
    // 2. U12 = L11^-1 A12 (L11 unit lower, w x w), one column per lane of
    // wave 0
    for (int e = tid; e < NB * NB; e += T) {
        const int c = e / NB, r = e % NB;
        sh.Lt[c][r] = (c < w && r < w) ? A[(int64_t)(k0 + c) * n + k0 + r] : 0.0f;
        sh.Ut[c][r] = (c < wj && r < w) ? A[(int64_t)(c0 + c) * n + k0 + r] : 0.0f;
    }
    __syncthreads();
    if (tid < wj) {
        // column tid of U12 stays in LDS (its row stride NB + 1 keeps the
        // 32 lanes on distinct banks); L11 reads are broadcasts
        float *u = sh.Ut[tid];
        for (int c = 0; c < w; ++c) {
            const float uc = u[c];
#pragma unroll 4
            for (int r = c + 1; r < w; ++r) u[r] = __builtin_fmaf(-sh.Lt[c][r], uc, u[r]);
        }
        for (int r = 0; r < w; ++r) A[(int64_t)(c0 + tid) * n + k0 + r] = u[r];
    }
    __syncthreads();
    // 3. A22 -= L21 U12 on the matrix cores
    lu_block_update<T>(A, n, k0 + w, c0, wj, k0, w, sh.Ut);
    // 4. lookahead: the next panel is this block, now up to date
    if (j == k + 1) {
        __syncthreads();
        lu_factor_panel<T>(A, n, c0, ipiv, info, sw + j, sh);
    }
}

// ------------------------------------- blocked Cholesky, n <= 1024 --------
// For the symmetric positive definite normal matrix BᵀB of the
// least-squares ('prescribed') system: A = L Lᵀ, right-looking in panels of
// NB columns with lookahead, one launch per panel (as the LU, without the
// per-column pivot search):
//   chol_first     factors panel 0;
//   chol_step(k)   one workgroup per column block j > k: rows >= j NB of
//                  the block -= L(:, panel k) L(block j rows, panel k)ᵀ on
//                  the matrix cores; block k + 1 keeps its updated rows in
//                  LDS and factors them as the next panel.
// A panel factor: wave 0 eliminates the NB x NB diagonal block and the 64 -
// NB rows below it in its lanes (row c's entries come by readlane, so no
// barrier per column); the other waves' rows solve x L11ᵀ = a by forward
// substitution against L11 in LDS, one row per thread.  The diagonal of L
// goes to dg.  A pivot d <= 0 (not positive definite in float32) sets
// info to its column + 1 and continues with 1.  Two elementwise kernels
// then rewrite L in the getrf form of the LU (chol_to_lu_*).
struct alignas(16) CholShared {
    float Ld[NB][NB + 4];  // L11 [row][column] (16-byte rows)
    float inv[NB];         // 1 / L11[c][c]
    float Ut[NB][NB + 1];  // [c][kk] = L(j0 + c, k0 + kk)
};

// rows: the panel's rows in LDS (row-major, stride NB + 1), or nullptr to
// read them from A.
template <int T>
__device__ __forceinline__ void chol_factor_panel(float *__restrict__ A, int n, int c0,
                                                  int32_t *__restrict__ info,
                                                  float *__restrict__ dg, CholShared &sh,
                                                  const float *__restrict__ rows)
{
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int w = min(NB, n - c0);
    const int row = c0 + tid;
    const bool live = row < n;
    float a[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j)
        a[j] = (live && j < w) ? (rows ? rows[tid * (NB + 1) + j] : A[(int64_t)(c0 + j) * n + row])
                               : 0.0f;
    if (wv == 0) {
        // Columns w .. NB - 1 of a narrow last panel are the identity (a
        // unit diagonal and zeros; their rows are past n), so every column
        // is eliminated the same way, without a branch: a branch target
        // inside the sequence makes the compiler wait for every memory
        // operation in flight at each column.
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j >= w && lane == j) a[j] = 1.0f;
        // branch-free: lanes at or above row c run the update with a zero
        // multiplier (a divergent branch would copy the row array at its
        // merge); 1 / lcc by the hardware reciprocal square root (the
        // factor is refined in float64 afterwards), kept by lane c
        int bad = -1;
        float invl = 1.0f;
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            const float d = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a[c]), c));
            const bool pos = d > 0.0f;
            if (!pos && bad < 0) bad = c;
            const float inv = pos ? __builtin_amdgcn_rsqf(d) : 1.0f;
            const float lcc = pos ? d * inv : 1.0f;
            // L(r, c) = a(r, c) / lcc; a(r, j) -= L(r, c) L(j, c) with
            // L(j, c) = a(c, j) / lcc by symmetry
            const bool below = lane > c;
            const float l = below ? a[c] * inv : 0.0f, li = l * inv;
            float rowc[NB];
#pragma unroll
            for (int j = c + 1; j < NB; ++j)
                rowc[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a[j]), c));
#pragma unroll
            for (int j = c + 1; j < NB; ++j) a[j] = __builtin_fmaf(-li, rowc[j], a[j]);
            a[c] = below ? l : (lane == c ? lcc : a[c]);
            invl = lane == c ? inv : invl;
            // column by column: interleaving columns chains each a[j]'s
            // readlane and update across columns, with wait states between
            __builtin_amdgcn_sched_barrier(0);
        }
        if (lane < NB) sh.inv[lane] = invl;
        if (lane < NB) {
#pragma unroll
            for (int q = 0; q < NB / 4; ++q)
                *reinterpret_cast<float4 *>(&sh.Ld[lane][4 * q]) =
                    make_float4(a[4 * q], a[4 * q + 1], a[4 * q + 2], a[4 * q + 3]);
        }
        if (lane == 0 && bad >= 0 && info[0] == 0) info[0] = c0 + bad + 1;
    }
    __syncthreads();
    if (wv > 0 && live) {
        // x L11ᵀ = a by forward substitution, L11 rows read 16 bytes at a time
#pragma unroll
        for (int c = 0; c < NB; ++c) {
            float s = a[c];
#pragma unroll
            for (int q = 0; 4 * q < c; ++q) {
                const float4 l4 = *reinterpret_cast<const float4 *>(&sh.Ld[c][4 * q]);
                s = __builtin_fmaf(-a[4 * q], l4.x, s);
                if (4 * q + 1 < c) s = __builtin_fmaf(-a[4 * q + 1], l4.y, s);
                if (4 * q + 2 < c) s = __builtin_fmaf(-a[4 * q + 2], l4.z, s);
                if (4 * q + 3 < c) s = __builtin_fmaf(-a[4 * q + 3], l4.w, s);
            }
            a[c] = s * sh.inv[c];
        }
    }
    // the lower part of the panel (row >= column)
    if (live) {
#pragma unroll
        for (int j = 0; j < NB; ++j)
            if (j < w && tid >= j) A[(int64_t)(c0 + j) * n + row] = a[j];
    }
    if (tid < w) dg[c0 + tid] = sh.Ld[tid][tid];
}

template <int T>
__global__ __launch_bounds__(T) void chol_first(float *__restrict__ A, int n,
                                                int32_t *__restrict__ info, float *__restrict__ dg)
{
    __shared__ CholShared sh;
    if (threadIdx.x == 0) info[0] = 0;
    __syncthreads();
    chol_factor_panel<T>(A, n, 0, info, dg, sh, nullptr);
}

template <int T>
__global__ __launch_bounds__(T) void chol_step(float *__restrict__ A, int n, int k,
                                               int32_t *__restrict__ info, float *__restrict__ dg)
{
    constexpr bool kRowsInLds = T <= 512;
    __shared__ CholShared sh;
    __shared__ float rows[kRowsInLds ? T * (NB + 1) : 1];
    const int j = k + 1 + blockIdx.x;
    const int tid = threadIdx.x;
    const int k0 = k * NB, w = min(NB, n - k0);
    const int c0 = j * NB, wj = min(NB, n - c0);
    // Q = L(block j rows, panel k)ᵀ, read along the rows of each column; its
    // loads and the update's operand loads are in flight together
    constexpr int QPER = (NB * NB + T - 1) / T;
    float qv[QPER];
#pragma unroll
    for (int it = 0; it < QPER; ++it) {
        const int e = tid + T * it, kk = e / NB, c = e % NB;
        qv[it] = (e < NB * NB && c < wj && kk < w) ? A[(int64_t)(k0 + kk) * n + c0 + c] : 0.0f;
    }
    const bool next = j == k + 1;
    float *dst = next && kRowsInLds ? rows : nullptr;
    lu_block_update<T>(A, n, c0, c0, wj, k0, w, sh.Ut, dst, [&]() {
#pragma unroll
        for (int it = 0; it < QPER; ++it) {
            const int e = tid + T * it;
            if (e < NB * NB) sh.Ut[e % NB][e / NB] = qv[it];
        }
        lds_barrier();
    });
    if (next) {
        __syncthreads();
        chol_factor_panel<T>(A, n, c0, info, dg, sh, dst);
    }
}

// The factor in LAPACK getrf form, for the inverse and the refinement:
// L Lᵀ = (L D⁻¹)(D Lᵀ), D = diag(L), pivots the identity.  One thread per
// pair (i, j), i >= j, of the lower triangle: it reads L(i, j) and writes
// both L(i, j) / D(j) below the diagonal and D(j) L(i, j) = U(j, i) above
// it (D(j)² on the diagonal), so no entry is read after another thread
// wrote it.  Lt (nullable): also the strictly lower part of L D⁻¹,
// row-major ldt x ldt (ldt >= n), zeros elsewhere (for spd_inverse); the
// threads then cover ldt x ldt.
__global__ void chol_to_lu(float *__restrict__ A, int n, const float *__restrict__ dg,
                           int32_t *__restrict__ ipiv, float *__restrict__ Lt, int ldt)
{
    const int m = Lt ? ldt : n;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)m * m) return;
    const int j = (int)(e / m), i = (int)(e % m);  // column j, row i
    if (Lt && (i <= j || i >= n)) Lt[(int64_t)i * ldt + j] = 0.0f;
    if (i < j || i >= n) return;
    const int64_t ea = (int64_t)j * n + i;
    const float d = dg[j];
    if (i == j) {
        A[ea] = d * d;
        ipiv[j] = j;
        return;
    }
    const float l = A[ea];
    A[ea] = l / d;
    A[(int64_t)i * n + j] = d * l;
    if (Lt) Lt[(int64_t)i * ldt + j] = l / d;
}

template <int T>
static void chol_launch(Lane *L, float *A, int n, int k, int nsteps, int32_t *info, float *dg)
{
    if (k < 0)
        chol_first<T><<<1, T, 0, L->stream>>>(A, n, info, dg);
    else if (nsteps - k - 1 > 0)
        chol_step<T><<<nsteps - k - 1, T, 0, L->stream>>>(A, n, k, info, dg);
}

// ------------------------------------------------- approximate inverse ----
// X = (P L U)^-1 in float32, column k from the right-hand side e_k (one
// workgroup per column; blocked substitution in LDS: a wave solves each
// NB x NB diagonal block, then all threads update the rows beyond it).
// X is stored row-major.  The refinement applies it as a matrix-vector
// product, so every sweep is two or three memory-bound GEMVs instead of a
// latency-bound triangular solve.
constexpr int SOLVE_T = 512;

__global__ __launch_bounds__(SOLVE_T) void lu_inverse_kernel(const float *__restrict__ LU, int n,
                                                             const int32_t *__restrict__ ipiv,
                                                             float *__restrict__ X)
{
    extern __shared__ float y[];  // n
    const int tid = threadIdx.x;
    const int col = blockIdx.x;
    for (int i = tid; i < n; i += SOLVE_T) y[i] = i == col ? 1.0f : 0.0f;
    __syncthreads();
    if (tid == 0)
        for (int k = 0; k < n; ++k) {
            const int p = ipiv[k];
            if (p != k) {
                const float t = y[k];
                y[k] = y[p];
                y[p] = t;
            }
        }
    __syncthreads();
    // forward: L unit lower
    for (int b0 = 0; b0 < n; b0 += NB) {
        const int w = min(NB, n - b0);
        if (tid < 64) {
            for (int c = 0; c < w; ++c) {
                const float xc = y[b0 + c];
                if (tid > c && tid < w)
                    y[b0 + tid] = __builtin_fmaf(-LU[(int64_t)(b0 + c) * n + b0 + tid], xc,
                                                 y[b0 + tid]);
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_s_waitcnt(0);
            }
        }
        __syncthreads();
        for (int i = b0 + w + tid; i < n; i += SOLVE_T) {
            float acc = y[i];
            for (int c = 0; c < w; ++c)
                acc = __builtin_fmaf(-LU[(int64_t)(b0 + c) * n + i], y[b0 + c], acc);
            y[i] = acc;
        }
        __syncthreads();
    }
    // backward: U upper, non-unit
    const int nblk = (n + NB - 1) / NB;
    for (int bk = nblk - 1; bk >= 0; --bk) {
        const int b0 = bk * NB, w = min(NB, n - b0);
        if (tid < 64) {
            for (int c = w - 1; c >= 0; --c) {
                if (tid == c) y[b0 + c] = y[b0 + c] / LU[(int64_t)(b0 + c) * n + b0 + c];
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_s_waitcnt(0);
                const float xc = y[b0 + c];
                if (tid < c)
                    y[b0 + tid] = __builtin_fmaf(-LU[(int64_t)(b0 + c) * n + b0 + tid], xc,
                                                 y[b0 + tid]);
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_s_waitcnt(0);
            }
        }
        __syncthreads();
        for (int i = tid; i < b0; i += SOLVE_T) {
            float acc = y[i];
            for (int c = 0; c < w; ++c)
                acc = __builtin_fmaf(-LU[(int64_t)(b0 + c) * n + i], y[b0 + c], acc);
            y[i] = acc;
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += SOLVE_T) X[(int64_t)i * n + col] = y[i];
}

// The same inverse for n <= 1024, 16 columns of X per workgroup of 16 waves.
// Column c of P E is the unit vector at the position the row swaps take
// row j0 + c to (one lane per column follows it through ipiv); then
// Y = L^-1 P E and X = U^-1 Y are solved in blocks of 16 rows kept in LDS.
// The update of a block by every block before it (after it, for U) is a
// 16 x 16 x K product on v_mfma_f32_16x16x4f32 with K split over the 16
// waves, whose partial tiles are added in wave order by 256 threads, one
// per tile entry; the 16 x 16 triangle is then solved by 16 lanes, one per
// column.  The factor entries of the next block are loaded while the
// current one is finished (the barriers order LDS only, so those loads
// stay in flight).
constexpr int INV_C = 16;
constexpr int INV_NMAX = 1024;
constexpr int INV_W = 16;                          // waves per workgroup
constexpr int INV_SMAX = INV_NMAX / 4 / INV_W;     // k-steps of 4 per wave and block

__global__ __launch_bounds__(64 * INV_W) void lu_inverse_mfma(const float *__restrict__ LU,
                                                              int n,
                                                              const int32_t *__restrict__ ipiv,
                                                              float *__restrict__ X)
{
    extern __shared__ float y[];  // y[r * INV_C + c], rows padded to 16
    __shared__ floatx4 part[INV_W][64];
    __shared__ float dg[16][17];  // the diagonal 16 x 16 block of the factors
    __shared__ int pos[INV_C];
    constexpr int T = 64 * INV_W;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int j0 = blockIdx.x * INV_C;
    const int nb = (n + 15) / 16, npad = 16 * nb;
    if (tid < INV_C) {
        int q = j0 + tid;
        if (q < n)
            for (int k = 0; k < n; ++k) {
                const int p = ipiv[k];
                if (q == k)
                    q = p;
                else if (q == p)
                    q = k;
            }
        pos[tid] = q;
    }
    __syncthreads();
    for (int e = tid; e < npad * INV_C; e += T) {
        const int r = e / INV_C, c = e % INV_C;
        y[e] = (j0 + c < n && r == pos[c]) ? 1.0f : 0.0f;
    }
    float av[INV_SMAX];
    float dgv = 0.0f;
    // factor entries of the block at rows rb for k in [ka, kb): this wave's
    // share of the products, and one entry of the diagonal block
    auto prefetch = [&](int rb, int ka, int kb) {
        const int steps = (kb - ka) / 4, per = (steps + INV_W - 1) / INV_W, s0 = wv * per;
        const int i = rb + li;
#pragma unroll
        for (int u = 0; u < INV_SMAX; ++u) {
            const int st = s0 + u, k = ka + 4 * st + lk;
            const bool ok = u < per && st < steps;
            av[u] = (ok && i < n && k < n) ? LU[(int64_t)k * n + i] : 0.0f;
        }
        if (tid < 256) {
            const int dr = tid >> 4, dc = tid & 15;
            dgv = (rb + dr < n && rb + dc < n) ? LU[(int64_t)(rb + dc) * n + rb + dr] : 0.0f;
        }
    };
    // this wave's partial tile, then (after a barrier) y[block] -= sum of
    // the partials in wave order
    auto products = [&](int ka, int kb) {
        const int steps = (kb - ka) / 4, per = (steps + INV_W - 1) / INV_W, s0 = wv * per;
        floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int u = 0; u < INV_SMAX; ++u) {
            const int st = s0 + u, k = ka + 4 * st + lk;
            if (u < per && st < steps)  // uniform across the wave
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[u], y[k * INV_C + li], acc, 0, 0,
                                                           0);
        }
        part[wv][lane] = acc;
    };
    auto subtract = [&](int rb, float dcur) {
        if (tid < 256) {
            const int r = tid >> 4, c = tid & 15;
            const int l2 = (r >> 2) * 16 + c, v = r & 3;
            float t = 0.0f;
#pragma unroll
            for (int q = 0; q < INV_W; ++q) t += part[q][l2][v];
            y[(rb + r) * INV_C + c] -= t;
            dg[r][c] = dcur;
        }
    };
    prefetch(0, 0, 0);
    __syncthreads();
    // forward: L unit lower
    for (int b = 0; b < nb; ++b) {
        const int rb = 16 * b;
        const float dcur = dgv;  // this block's diagonal entry (dgv is refilled next)
        products(0, rb);
        if (b + 1 < nb)
            prefetch(rb + 16, 0, rb + 16);
        lds_barrier();
        subtract(rb, dcur);
        lds_barrier();
        if (tid < INV_C) {
            float col[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) col[r] = y[(rb + r) * INV_C + tid];
#pragma unroll
            for (int r = 0; r < 16; ++r)
#pragma unroll
                for (int rr = 0; rr < 16; ++rr)
                    if (rr > r) col[rr] = __builtin_fmaf(-dg[rr][r], col[r], col[rr]);
#pragma unroll
            for (int r = 0; r < 16; ++r) y[(rb + r) * INV_C + tid] = col[r];
        }
        lds_barrier();
    }
    // backward: U upper, non-unit
    prefetch(npad - 16, npad, npad);
    for (int b = nb - 1; b >= 0; --b) {
        const int rb = 16 * b;
        const float dcur = dgv;
        products(rb + 16, npad);
        if (b > 0) prefetch(rb - 16, rb, npad);
        lds_barrier();
        subtract(rb, dcur);
        lds_barrier();
        if (tid < INV_C) {
            float col[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) col[r] = y[(rb + r) * INV_C + tid];
#pragma unroll
            for (int r = 15; r >= 0; --r) {
                if (rb + r < n) col[r] = col[r] / dg[r][r];
#pragma unroll
                for (int rr = 0; rr < 16; ++rr)
                    if (rr < r) col[rr] = __builtin_fmaf(-dg[rr][r], col[r], col[rr]);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) y[(rb + r) * INV_C + tid] = col[r];
        }
        lds_barrier();
    }
    for (int e = tid; e < n * INV_C; e += T) {
        const int r = e / INV_C, c = e % INV_C;
        if (j0 + c < n) X[(int64_t)r * n + j0 + c] = y[e];
    }
}

// ------------------------------ inverse of the Cholesky-factored matrix ---
// The Cholesky factors in getrf form are N = L' D² L'ᵀ (L' unit lower, D²
// the diagonal of U), so N⁻¹ = W'ᵀ D⁻² W' with W' = L'⁻¹.  W' is formed by
// doubling rather than by n column solves: the 32 x 32 diagonal blocks are
// inverted first (one wave each), then for h = 32, 64, ... each pair of
// inverted h-blocks of [L11 0; L21 L22] is completed by
// W21 = -W22 (L21 W11), two products on the matrix cores per level; X =
// W'ᵀ D⁻² W' follows in 32 x 32 tiles over the whole matrix.  The work
// arrays are npad x npad (npad = n rounded up to 32), and the rows and
// columns past n are the identity in W', which adds nothing to X.  The
// products read their operands 16 bytes per lane: lanes 16 q .. 16 q + 15
// take k = k0 + 4 q + s at step s of a 16-wide chunk of the sum (an order
// of the terms, the same for both operands), so A must be row-major and B
// column-major.  The k ranges skip the blocks above the diagonal of W',
// which are never written.
constexpr int SI_B = 32;

// Block b of W' = L'⁻¹ (a 32 x 32 unit lower block): lane j forms column j
// by forward substitution against the block of L' (rows of Lt in LDS,
// zeros on and above the diagonal), written column-major to Wc and
// row-major to Wr, zeros above the diagonal included.
__global__ __launch_bounds__(64) void spd_inv_diag(const float *__restrict__ Lt, int npad,
                                                   float *__restrict__ Wc, float *__restrict__ Wr)
{
    __shared__ float4 Ls[SI_B][SI_B / 4];
    const int lane = threadIdx.x, r0 = SI_B * blockIdx.x;
    for (int e = lane; e < SI_B * SI_B / 4; e += 64) {
        const int r = e / (SI_B / 4), q = e % (SI_B / 4);
        Ls[r][q] = *reinterpret_cast<const float4 *>(Lt + (int64_t)(r0 + r) * npad + r0 + 4 * q);
    }
    __syncthreads();
    if (lane >= SI_B) return;
    const int j = lane;
    float w[SI_B];  // entries above j come out 0: their sums have only zero terms
#pragma unroll
    for (int i = 0; i < SI_B; ++i) {
        float s = i == j ? 1.0f : 0.0f;
#pragma unroll
        for (int q = 0; 4 * q < i; ++q) {
            const float4 l4 = Ls[i][q];
            s = __builtin_fmaf(-l4.x, w[4 * q], s);
            if (4 * q + 1 < i) s = __builtin_fmaf(-l4.y, w[4 * q + 1], s);
            if (4 * q + 2 < i) s = __builtin_fmaf(-l4.z, w[4 * q + 2], s);
            if (4 * q + 3 < i) s = __builtin_fmaf(-l4.w, w[4 * q + 3], s);
        }
        w[i] = s;
    }
    float *wc = Wc + (int64_t)(r0 + j) * npad + r0;
#pragma unroll
    for (int q = 0; q < SI_B / 4; ++q)
        *reinterpret_cast<float4 *>(wc + 4 * q) =
            make_float4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
#pragma unroll
    for (int i = 0; i < SI_B; ++i) Wr[(int64_t)(r0 + i) * npad + r0 + j] = w[i];
}

// 16 x 16 tile of A B on the matrix cores over k in [ka, kb) (multiples of
// 16): a points at row li of A (row-major), b at column li of B
// (column-major), both at k = 0; the result lane holds D(4 q + v, li).
__device__ __forceinline__ floatx4 spd_tile(const float *__restrict__ a, const float *__restrict__ b,
                                            int ka, int kb, int q)
{
    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    int k = ka;
    // four 16-wide chunks' operands in flight before their products
    for (; k + 48 < kb; k += 64) {
        float4 a4[4], b4[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a4[u] = *reinterpret_cast<const float4 *>(a + k + 16 * u + 4 * q);
            b4[u] = *reinterpret_cast<const float4 *>(b + k + 16 * u + 4 * q);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[u].x, b4[u].x, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[u].y, b4[u].y, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[u].z, b4[u].z, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[u].w, b4[u].w, acc, 0, 0, 0);
        }
    }
    for (; k < kb; k += 16) {
        const float4 a4 = *reinterpret_cast<const float4 *>(a + k + 4 * q);
        const float4 b4 = *reinterpret_cast<const float4 *>(b + k + 4 * q);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.x, b4.x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.y, b4.y, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.z, b4.z, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.w, b4.w, acc, 0, 0, 0);
    }
    return acc;
}

// Level h of the doubling for the pair of blocks at r0 = 2 h blockIdx.y
// (the second block has h2 = min(h, npad - r0 - h) rows); a workgroup
// computes a 32 x 32 tile (ti, tj) of the h2 x h result, a wave 16 x 16 of
// it.  STAGE 0: Wt = L21 W11, k from the tile's first column (W11 is lower
// triangular), stored where W21 goes; STAGE 1: W21 = -W22 Wt, k up to the
// tile's last row (W22 is lower triangular).
template <int STAGE>
__global__ __launch_bounds__(256) void spd_inv_level(const float *__restrict__ Lt,
                                                     float *__restrict__ Wc,
                                                     float *__restrict__ Wr,
                                                     float *__restrict__ Wt, int npad, int h)
{
    const int r0 = 2 * h * blockIdx.y, h2 = min(h, npad - r0 - h);
    const int tiles = h / SI_B, ti = blockIdx.x / tiles, tj = blockIdx.x % tiles;
    if (h2 <= 0 || SI_B * ti >= h2) return;  // uniform across the workgroup
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, q = lane >> 4;
    const int i0 = SI_B * ti + 16 * (wv & 1), j0 = SI_B * tj + 16 * (wv >> 1);
    const int row = r0 + h + i0, col = r0 + j0;  // of this wave's tile in W'
    floatx4 acc;
    if (STAGE == 0) {
        acc = spd_tile(Lt + (int64_t)(row + li) * npad + r0, Wc + (int64_t)(col + li) * npad + r0,
                       SI_B * tj, h, q);
        *reinterpret_cast<float4 *>(Wt + (int64_t)(col + li) * npad + row + 4 * q) =
            make_float4(acc[0], acc[1], acc[2], acc[3]);
    } else {
        acc = spd_tile(Wr + (int64_t)(row + li) * npad + r0 + h,
                       Wt + (int64_t)(col + li) * npad + r0 + h, 0, SI_B * (ti + 1), q);
        *reinterpret_cast<float4 *>(Wc + (int64_t)(col + li) * npad + row + 4 * q) =
            make_float4(-acc[0], -acc[1], -acc[2], -acc[3]);
#pragma unroll
        for (int v = 0; v < 4; ++v) Wr[(int64_t)(row + 4 * q + v) * npad + col + li] = -acc[v];
    }
}

// X = W'ᵀ D⁻² W' (n x n, row-major), tile (blockIdx.x, blockIdx.y) with k
// from the larger of its row and column offsets (W' is lower triangular).
// D² is the diagonal of the Cholesky factor squared.  X is symmetric, so a
// lane's D(i, j) for rows i = 4 q + v of column j goes to the row-major
// place of (j, i): 16 bytes along row j.
__global__ __launch_bounds__(256) void spd_inv_x(const float *__restrict__ Wc,
                                                 const float *__restrict__ dg, int n, int npad,
                                                 float *__restrict__ X)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, q = lane >> 4;
    const int i0 = SI_B * blockIdx.x + 16 * (wv & 1), j0 = SI_B * blockIdx.y + 16 * (wv >> 1);
    const float *a = Wc + (int64_t)(i0 + li) * npad, *b = Wc + (int64_t)(j0 + li) * npad;
    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    // D⁻² by the hardware reciprocal (X preconditions a float64 refinement);
    // two 16-wide chunks' operands in flight before their products
    auto chunk = [&](int k, float4 &a4, float4 &b4, float s[4]) {
        const int kq = k + 4 * q;
        a4 = *reinterpret_cast<const float4 *>(a + kq);
        b4 = *reinterpret_cast<const float4 *>(b + kq);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const float d = kq + t < n ? dg[kq + t] : 1.0f;
            s[t] = kq + t < n ? __builtin_amdgcn_rcpf(d * d) : 0.0f;
        }
    };
    auto products = [&](const float4 &a4, const float4 &b4, const float s[4]) {
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.x * s[0], b4.x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.y * s[1], b4.y, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.z * s[2], b4.z, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4.w * s[3], b4.w, acc, 0, 0, 0);
    };
    int k = SI_B * max(blockIdx.x, blockIdx.y);
    for (; k + 16 < npad; k += 32) {
        float4 a0, b0, a1, b1;
        float s0[4], s1[4];
        chunk(k, a0, b0, s0);
        chunk(k + 16, a1, b1, s1);
        products(a0, b0, s0);
        products(a1, b1, s1);
    }
    for (; k < npad; k += 16) {
        float4 a0, b0;
        float s0[4];
        chunk(k, a0, b0, s0);
        products(a0, b0, s0);
    }
    const int j = j0 + li, i = i0 + 4 * q;
    if (j >= n) return;
    float *x = X + (int64_t)j * n + i;
    if ((n & 3) == 0 && i + 3 < n) {
        *reinterpret_cast<float4 *>(x) = make_float4(acc[0], acc[1], acc[2], acc[3]);
    } else {
#pragma unroll
        for (int v = 0; v < 4; ++v)
            if (i + v < n) x[v] = acc[v];
    }
}

// Xinv = N⁻¹ from the Cholesky factors of f (Lt, dg filled by the
// factorisation).
static int spd_inverse(Lane *L, DenseFactors *f)
{
    const int n = (int)f->n, npad = SI_B * (int)div_up(f->n, SI_B), nb = npad / SI_B;
    spd_inv_diag<<<nb, 64, 0, L->stream>>>(f->Lt, npad, f->Wc, f->Wr);
    for (int h = SI_B; h < npad; h *= 2) {
        const dim3 grid((h / SI_B) * (h / SI_B), (unsigned)div_up(npad, 2 * h));
        spd_inv_level<0><<<grid, 256, 0, L->stream>>>(f->Lt, f->Wc, f->Wr, f->Wt, npad, h);
        spd_inv_level<1><<<grid, 256, 0, L->stream>>>(f->Lt, f->Wc, f->Wr, f->Wt, npad, h);
    }
    spd_inv_x<<<dim3(nb, nb), 256, 0, L->stream>>>(f->Wc, f->dg, n, npad, f->Xinv);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

// ------------------------------------------------- refinement kernels -----
// Convergence state of one refinement, kept on the device so that sweeps
// run back to back; every kernel of a sweep returns at once when done.
struct RefineState {
    double last_dx;
    int32_t done, iterations, converged, pad;
    // one sweep's max |d| and max |x| as the bit patterns of non-negative
    // doubles (their unsigned order is the numeric order, NaN above
    // infinity), and the workgroups of the sweep that have finished
    unsigned long long md_bits, mx_bits;
    uint32_t finished, pad2;
};

__device__ __forceinline__ double wave_sum_f64(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;  // fixed butterfly: every lane holds the same sum
}

// warm: x holds a starting guess (kept); otherwise x = 0
__global__ void refine_init(double *__restrict__ x, int64_t n, int warm,
                            RefineState *__restrict__ st)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !warm) x[i] = 0.0;
    if (i == 0) {
        st->last_dx = INFINITY;
        st->done = 0;
        st->iterations = 0;
        st->converged = 0;
        st->md_bits = 0ull;
        st->mx_bits = 0ull;
        st->finished = 0u;
    }
}

// r = c - B x  (B row-major rows x n, float64), one wave per row
__global__ __launch_bounds__(256) void resid_rows(const double *__restrict__ Br, int64_t rows,
                                                  int64_t n, const double *__restrict__ x,
                                                  const double *__restrict__ c,
                                                  double *__restrict__ r,
                                                  const RefineState *__restrict__ st)
{
    if (st->done) return;
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= rows) return;
    const double *row = Br + i * n;
    double s = 0.0;
    for (int64_t j = lane; j < n; j += 64) s = __builtin_fma(row[j], x[j], s);
    s = wave_sum_f64(s);
    if (lane == 0) r[i] = c[i] - s;
}

// g = B^T r  (B column-major rows x n, float64), one wave per column
__global__ __launch_bounds__(256) void tapply_cols(const double *__restrict__ Bc, int64_t rows,
                                                   int64_t n, const double *__restrict__ r,
                                                   double *__restrict__ g,
                                                   const RefineState *__restrict__ st)
{
    if (st->done) return;
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= n) return;
    const double *col = Bc + j * rows;
    double s = 0.0;
    for (int64_t i = lane; i < rows; i += 64) s = __builtin_fma(col[i], r[i], s);
    s = wave_sum_f64(s);
    if (lane == 0) g[j] = s;
}

// One sweep's correction and update: d = X v (X float32 row-major n x n,
// products and sums in float64, one wave per row), x += d, then the
// stopping rule: the correction has reached float64 resolution (|d| <= 4
// eps |x|), or it stopped shrinking (after the third sweep, |d| >=
// |d_prev| / 2: converged if |d| <= 1e-10 |x|), or a non-finite value, or
// max_it sweeps.  Each workgroup folds its max |d| and max |x| into the
// state; the last one to finish applies the rule and clears the folds for
// the next sweep (no workgroup waits for another).
__global__ __launch_bounds__(256) void xapply_update(const float *__restrict__ X, int64_t n,
                                                     const double *__restrict__ v,
                                                     double *__restrict__ x,
                                                     RefineState *__restrict__ st, int max_it)
{
    if (st->done) return;
    __shared__ unsigned long long smd[4], smx[4];
    __shared__ int last;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t i = (int64_t)blockIdx.x * 4 + wv;
    unsigned long long bd = 0ull, bx = 0ull;
    if (i < n) {
        const float *row = X + i * n;
        double s = 0.0;
        for (int64_t j = lane; j < n; j += 64) s = __builtin_fma((double)row[j], v[j], s);
        s = wave_sum_f64(s);
        const double xn = x[i] + s;
        if (lane == 0) x[i] = xn;
        bd = (unsigned long long)__double_as_longlong(fabs(s));
        bx = (unsigned long long)__double_as_longlong(fabs(xn));
    }
    if (lane == 0) {
        smd[wv] = bd;
        smx[wv] = bx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long md = smd[0], mx = smx[0];
        for (int k = 1; k < 4; ++k) {
            md = smd[k] > md ? smd[k] : md;
            mx = smx[k] > mx ? smx[k] : mx;
        }
        atomicMax(&st->md_bits, md);
        atomicMax(&st->mx_bits, mx);
        __threadfence();
        last = atomicAdd(&st->finished, 1u) == gridDim.x - 1u;
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    __threadfence();
    const double dmax = __longlong_as_double((long long)atomicMax(&st->md_bits, 0ull));
    const double xmax = __longlong_as_double((long long)atomicMax(&st->mx_bits, 0ull));
    const int it = st->iterations;
    const double eps = 2.220446049250313e-16;
    int stop = 0, conv = 0;
    if (!isfinite(dmax) || !isfinite(xmax)) {
        stop = 1;
    } else if (dmax <= 4.0 * eps * xmax) {
        stop = conv = 1;
    } else if (it > 2 && dmax >= 0.5 * st->last_dx) {
        stop = 1;
        conv = dmax <= 1e-10 * xmax;
    }
    st->iterations = it + 1;
    st->last_dx = dmax;
    if (stop || it + 1 >= max_it) {
        st->done = 1;
        st->converged = conv;
    }
    st->md_bits = 0ull;
    st->mx_bits = 0ull;
    st->finished = 0u;
}

// ------------------------------------------------------------- host ------
size_t lu_work_bytes(int n) { return sizeof(LuSwaps) * (size_t)std::max(1, (n + NB - 1) / NB); }

// Workgroup size for a panel of `rows` live rows: one row per thread, as
// few waves as cover them (the panel loop is bound by the waves' vector
// instruction issue, so idle waves would cost as much as busy ones).
static int lu_threads(int rows)
{
    int t = 64;
    while (t < rows) t *= 2;
    return t;
}

template <int T>
static void lu_launch_first(Lane *L, float *A32, int n, int32_t *ipiv, int32_t *info,
                            LuSwaps *sw)
{
    lu_first<T><<<1, T, 0, L->stream>>>(A32, n, ipiv, info, sw);
}

template <int T>
static void lu_launch_step(Lane *L, float *A32, int n, int k, int nsteps, int32_t *ipiv,
                           int32_t *info, LuSwaps *sw)
{
    lu_step<T><<<(unsigned)nsteps, T, 0, L->stream>>>(A32, n, k, ipiv, info, sw);
}

int lu_factor_f32(Lane *L, float *A32, int n, int32_t *ipiv, int32_t *info, void *work)
{
    if (n <= 0) return PHF_OK;
    if (n <= LU_TMAX) {
        const int nsteps = (n + NB - 1) / NB;
        LuSwaps *sw = reinterpret_cast<LuSwaps *>(work);
        if (!sw) PHF_ALLOC(sw, L->arena, LuSwaps, (size_t)nsteps);
        for (int k = -1; k < nsteps; ++k) {
            // launch k + 1 factors panel k + 1 (rows from (k + 1) NB on);
            // k = -1 is the first panel alone
            const int t = lu_threads(std::max(1, n - (k + 1) * NB));
            switch (t) {
            case 64:
                k < 0 ? lu_launch_first<64>(L, A32, n, ipiv, info, sw)
                      : lu_launch_step<64>(L, A32, n, k, nsteps, ipiv, info, sw);
                break;
            case 128:
                k < 0 ? lu_launch_first<128>(L, A32, n, ipiv, info, sw)
                      : lu_launch_step<128>(L, A32, n, k, nsteps, ipiv, info, sw);
                break;
            case 256:
                k < 0 ? lu_launch_first<256>(L, A32, n, ipiv, info, sw)
                      : lu_launch_step<256>(L, A32, n, k, nsteps, ipiv, info, sw);
                break;
            case 512:
                k < 0 ? lu_launch_first<512>(L, A32, n, ipiv, info, sw)
                      : lu_launch_step<512>(L, A32, n, k, nsteps, ipiv, info, sw);
                break;
            default:
                k < 0 ? lu_launch_first<1024>(L, A32, n, ipiv, info, sw)
                      : lu_launch_step<1024>(L, A32, n, k, nsteps, ipiv, info, sw);
                break;
            }
        }
        PHF_HIP(hipGetLastError());
        return PHF_OK;
    }
    // larger systems: one kernel per panel stage
    PHF_HIP(hipMemsetAsync(info, 0, sizeof(int32_t), L->stream));
    const size_t lds_max = Context::get().props().sharedMemPerBlock;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int rows = n - k0;
        const size_t lds = sizeof(float) * (size_t)NB * rows;
        const int use_lds = lds + 4096 <= lds_max ? 1 : 0;
        lu_panel<<<1, 256, use_lds ? lds : 0, L->stream>>>(A32, n, k0, ipiv, info, use_lds);
        lu_swap_trsm<<<div_up(n, 256), 256, 0, L->stream>>>(A32, n, k0, ipiv);
        const int m = n - k0 - NB;
        if (m > 0) {
            const int t = (m + 15) / 16;
            lu_update_mfma<<<t * t, 64, 0, L->stream>>>(A32, n, k0);
        }
    }
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

int chol_factor_f32(Lane *L, float *A32, int n, int32_t *ipiv, int32_t *info, float *dg,
                    float *Lt, int ldt)
{
    if (n <= 0) return PHF_OK;
    PHF_REQUIRE(n <= LU_TMAX, "Cholesky factorisation of n = %d (at most %d)", n, LU_TMAX);
    const int nsteps = (n + NB - 1) / NB;
    for (int k = -1; k + 1 < nsteps; ++k) {
        // launch k + 1 factors panel k + 1 (rows from (k + 1) NB on)
        switch (lu_threads(std::max(1, n - (k + 1) * NB))) {
        case 64: chol_launch<64>(L, A32, n, k, nsteps, info, dg); break;
        case 128: chol_launch<128>(L, A32, n, k, nsteps, info, dg); break;
        case 256: chol_launch<256>(L, A32, n, k, nsteps, info, dg); break;
        case 512: chol_launch<512>(L, A32, n, k, nsteps, info, dg); break;
        default: chol_launch<1024>(L, A32, n, k, nsteps, info, dg); break;
        }
    }
    PHF_REQUIRE(!Lt || ldt >= n, "Cholesky: row stride %d of the unit lower copy below n = %d",
                ldt, n);
    const int m = Lt ? ldt : n;
    chol_to_lu<<<div_up((int64_t)m * m, 256), 256, 0, L->stream>>>(A32, n, dg, ipiv, Lt, ldt);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

__global__ void to_f32(const double *__restrict__ a, int64_t n, float *__restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

static int dense_inverse(Lane *L, DenseFactors *f)
{
    const int64_t n = f->n;
    const size_t inv_lds = sizeof(float) * INV_C * 16 * div_up(n, 16);
    if (n <= INV_NMAX && inv_lds + 8192 <= Context::get().props().sharedMemPerBlock)
        lu_inverse_mfma<<<div_up(n, INV_C), 64 * INV_W, inv_lds, L->stream>>>(
            f->LU, (int)n, f->ipiv, f->Xinv);
    else
        lu_inverse_kernel<<<(unsigned)n, SOLVE_T, sizeof(float) * (size_t)n, L->stream>>>(
            f->LU, (int)n, f->ipiv, f->Xinv);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

int dense_factor_enqueue(Lane *L, const double *B, int64_t rows, int64_t n, DenseFactors *f,
                         hipEvent_t *marks)
{
    const bool normal = rows > n;
    if (!(f->LU && f->rows == rows && f->n == n)) {
        f->rows = rows;
        f->n = n;
        PHF_ALLOC(f->Bc, L->arena, double, (size_t)(rows * n));
        PHF_ALLOC(f->LU, L->arena, float, (size_t)(n * n));
        PHF_ALLOC(f->ipiv, L->arena, int32_t, (size_t)n);
        PHF_ALLOC(f->info, L->arena, int32_t, 1);
        PHF_ALLOC(f->Xinv, L->arena, float, (size_t)(n * n));
        PHF_ALLOC(f->dg, L->arena, float, (size_t)n);
        if (normal) PHF_ALLOC(f->N64, L->arena, double, (size_t)(n * n));
        if (normal && n <= LU_TMAX) {
            const size_t np = (size_t)(SI_B * div_up(n, SI_B));
            PHF_ALLOC(f->Lt, L->arena, float, np * np);
            PHF_ALLOC(f->Wc, L->arena, float, np * np);
            PHF_ALLOC(f->Wr, L->arena, float, np * np);
            PHF_ALLOC(f->Wt, L->arena, float, np * np);
        }
        if (n <= LU_TMAX) {
            char *w;
            PHF_ALLOC(w, L->arena, char, lu_work_bytes((int)n));
            f->lu_work = w;
        }
    }
    to_colmajor<<<div_up(rows * n, 256), 256, 0, L->stream>>>(B, rows, n, f->Bc,
                                                              normal ? nullptr : f->LU);
    if (normal) {
        const dim3 grid(div_up(n, NM_T), div_up(n, NM_T));
        normal_matrix<<<grid, 256, 0, L->stream>>>(f->Bc, rows, n, f->N64, f->LU);
    }
    if (marks) PHF_HIP(hipEventRecord(marks[0], L->stream));
    // the normal matrix is symmetric positive definite: Cholesky (no pivot
    // search); B itself: LU with partial pivoting
    const bool chol = normal && n <= LU_TMAX;
    if (chol)
        PHF_TRY(chol_factor_f32(L, f->LU, (int)n, f->ipiv, f->info, f->dg, f->Lt,
                                SI_B * (int)div_up(n, SI_B)));
    else
        PHF_TRY(lu_factor_f32(L, f->LU, (int)n, f->ipiv, f->info, f->lu_work));
    if (marks) PHF_HIP(hipEventRecord(marks[1], L->stream));
    if (chol)
        PHF_TRY(spd_inverse(L, f));
    else
        PHF_TRY(dense_inverse(L, f));
    if (marks) PHF_HIP(hipEventRecord(marks[2], L->stream));
    f->Br = B;
    f->cholesky = chol;
    return PHF_OK;
}

int dense_factor_finish(Lane *L, DenseFactors *f, int32_t hinfo)
{
    const int64_t n = f->n;
    if (f->cholesky && hinfo != 0) {
        // not positive definite in float32: the pivoted LU of the same matrix
        to_f32<<<div_up(n * n, 256), 256, 0, L->stream>>>(f->N64, n * n, f->LU);
        PHF_TRY(lu_factor_f32(L, f->LU, (int)n, f->ipiv, f->info, f->lu_work));
        PHF_TRY(dense_inverse(L, f));
        f->cholesky = false;
        PHF_HIP(hipMemcpyAsync(&hinfo, f->info, sizeof(int32_t), hipMemcpyDeviceToHost,
                               L->stream));
        PHF_HIP(hipStreamSynchronize(L->stream));
    }
    f->singular_at = hinfo;
    if (hinfo != 0) {
        set_error("matrix is singular in float32 (zero pivot at column %d)", hinfo - 1);
        return PHF_ERR_INVALID;
    }
    return PHF_OK;
}

int dense_factor_device(Lane *L, const double *B, int64_t rows, int64_t n, DenseFactors *f,
                        hipEvent_t *marks)
{
    PHF_TRY(dense_factor_enqueue(L, B, rows, n, f, marks));
    int32_t hinfo = 0;
    PHF_HIP(hipMemcpyAsync(&hinfo, f->info, sizeof(int32_t), hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    return dense_factor_finish(L, f, hinfo);
}

int dense_refine_device(Lane *L, const DenseFactors &f, const double *c, double *x,
                        SolveStats *st, bool warm, int first_burst)
{
    const int64_t rows = f.rows, n = f.n;
    const bool normal = rows > n;
    double *r, *g;
    RefineState *dst;
    PHF_ALLOC(r, L->arena, double, (size_t)rows);
    PHF_ALLOC(g, L->arena, double, (size_t)n);
    PHF_ALLOC(dst, L->arena, RefineState, 1);
    refine_init<<<div_up(std::max<int64_t>(n, 1), 256), 256, 0, L->stream>>>(x, n, warm ? 1 : 0,
                                                                          dst);
    const int max_it = 60;
    RefineState hs{};
    for (int launched = 0, burst = std::max(1, first_burst); launched < max_it;
         launched += burst, burst = 6) {
        for (int b = 0; b < burst && launched + b < max_it; ++b) {
            resid_rows<<<div_up(rows, 4), 256, 0, L->stream>>>(f.Br, rows, n, x, c, r, dst);
            const double *rhs = r;
            if (normal) {
                tapply_cols<<<div_up(n, 4), 256, 0, L->stream>>>(f.Bc, rows, n, r, g, dst);
                rhs = g;
            }
            xapply_update<<<div_up(n, 4), 256, 0, L->stream>>>(f.Xinv, n, rhs, x, dst, max_it);
        }
        PHF_HIP(hipGetLastError());
        PHF_HIP(hipMemcpyAsync(&hs, dst, sizeof(hs), hipMemcpyDeviceToHost, L->stream));
        PHF_HIP(hipStreamSynchronize(L->stream));
        if (hs.done) break;
    }
    st->iterations = hs.iterations;
    st->converged = hs.converged != 0;
    return PHF_OK;
}

int dense_solve_device(Lane *L, const double *B, int64_t rows, int64_t n, const double *c,
                       double *x, SolveStats *st)
{
    hipEvent_t e0 = L->ev[0], e1 = L->ev[1], e2 = L->ev[2];
    PHF_HIP(hipEventRecord(e0, L->stream));
    DenseFactors f;
    int rc = dense_factor_device(L, B, rows, n, &f);
    st->singular_at = f.singular_at;
    if (rc != PHF_OK) return rc;
    PHF_HIP(hipEventRecord(e1, L->stream));
    PHF_TRY(dense_refine_device(L, f, c, x, st));
    PHF_HIP(hipEventRecord(e2, L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    float ms_lu = 0, ms_ref = 0;
    (void)hipEventElapsedTime(&ms_lu, e0, e1);
    (void)hipEventElapsedTime(&ms_ref, e1, e2);
    st->lu_ms = ms_lu;
    st->refine_ms = ms_ref;
    return PHF_OK;
}

// ---------------------------------------------- grouped panel far field --
// Panels are grouped 16 at a time (consecutive indices).  In complex form a
// panel's velocity is conj(v) = i gamma e^{-i alpha} / (2 pi) log((z - z1) /
// (z - z2)), z1/z2 its ends (the per-panel formula above, rearranged), so a
// group's velocity beyond 3 R_g of its centre z_g (R_g: farthest panel end)
// is the multipole series conj(v) = (1/R_g) sum_k b_k t^(k+1), t = R_g /
// (z - z_g), with b_k = sum_p i gamma_p e^{-i alpha_p} / (2 pi) R_g
// (zh2^(k+1) - zh1^(k+1)) / (k + 1), zh = (z_end - z_g) / R_g.  |t| <= 1/3:
// 16 terms leave < 3e-8 of the group's velocity (float32 path), 30 terms
// < 1e-14 (float64 path).  Nearer groups are summed panel by panel.
// Targets are evaluated in Morton order (16-bit keys over their bounding
// box) so that a wave's targets take the same branch for almost every
// group.  Each target has PI_LPT lanes: lane s sums the groups s, s +
// PI_LPT, ... in index order and the lanes' sums are added by a fixed
// butterfly (one lane per target left the SIMDs latency-bound: each
// group's series is a serial chain, and the VALU issued 29% of the time).
constexpr int PG_SIZE = 16;    // panels per group
constexpr int PG_TERMS = 30;   // series terms kept (float64; float32 uses 16)
constexpr int PG_CHUNK = 64;   // group headers and series staged per round
constexpr int PI_LPT = 4;      // lanes per target in the grouped evaluation

struct PanelG {
    double zx, zy;     // centre
    double far2;       // (3 R_g)^2
    double R, inv_R;   // R_g, 1 / R_g
    int32_t first, count, pad0, pad1;
};

__global__ __launch_bounds__(64) void panel_group_coeffs(
    int64_t M, const double *__restrict__ gamma, const double *__restrict__ xs,
    const double *__restrict__ ys, const double *__restrict__ xe, const double *__restrict__ ye,
    const double *__restrict__ cosA, const double *__restrict__ sinA, PanelG *__restrict__ grp,
    double2 *__restrict__ coef)
{
    __shared__ double pd[PG_SIZE][7];  // xs, ys, xe, ye, gamma, cos, sin
    __shared__ double sc[3];
    const int g = blockIdx.x, k = threadIdx.x;
    const int64_t p0 = (int64_t)g * PG_SIZE;
    const int cnt = (int)min((int64_t)PG_SIZE, M - p0);
    if (k < cnt) {
        const int64_t p = p0 + k;
        pd[k][0] = xs[p];
        pd[k][1] = ys[p];
        pd[k][2] = xe[p];
        pd[k][3] = ye[p];
        pd[k][4] = gamma[p];
        pd[k][5] = cosA[p];
        pd[k][6] = sinA[p];
    }
    __syncthreads();
    if (k == 0) {
        double cx = 0.0, cy = 0.0;
        for (int q = 0; q < cnt; ++q) {
            cx += 0.5 * (pd[q][0] + pd[q][2]);
            cy += 0.5 * (pd[q][1] + pd[q][3]);
        }
        cx /= cnt;
        cy /= cnt;
        double r2 = 0.0;
        for (int q = 0; q < cnt; ++q) {
            const double a1 = pd[q][0] - cx, b1 = pd[q][1] - cy;
            const double a2 = pd[q][2] - cx, b2 = pd[q][3] - cy;
            r2 = fmax(r2, fmax(a1 * a1 + b1 * b1, a2 * a2 + b2 * b2));
        }
        const double R = sqrt(r2);
        sc[0] = cx;
        sc[1] = cy;
        sc[2] = R;
        PanelG pg;
        pg.zx = cx;
        pg.zy = cy;
        pg.far2 = 9.0 * r2;
        pg.R = R;
        pg.inv_R = 1.0 / R;
        pg.first = (int32_t)p0;
        pg.count = cnt;
        pg.pad0 = pg.pad1 = 0;
        grp[g] = pg;
    }
    __syncthreads();
    if (k >= PG_TERMS) return;
    const double cx = sc[0], cy = sc[1], R = sc[2], iR = 1.0 / R;
    double bx = 0.0, by = 0.0;
    for (int q = 0; q < cnt; ++q) {
        // zh^(k+1) for both ends, zh = (end - z_g) / R_g
        const double u1 = (pd[q][0] - cx) * iR, v1 = (pd[q][1] - cy) * iR;
        const double u2 = (pd[q][2] - cx) * iR, v2 = (pd[q][3] - cy) * iR;
        double p1r = u1, p1i = v1, p2r = u2, p2i = v2;
        for (int e = 0; e < k; ++e) {
            const double t1 = p1r * u1 - p1i * v1;
            p1i = p1r * v1 + p1i * u1;
            p1r = t1;
            const double t2 = p2r * u2 - p2i * v2;
            p2i = p2r * v2 + p2i * u2;
            p2r = t2;
        }
        const double dr = (p2r - p1r) / (k + 1), di = (p2i - p1i) / (k + 1);
        // f = i gamma (c - i s) / (2 pi) R = gamma R / (2 pi) (s + i c)
        const double f = pd[q][4] * R / (2.0 * M_PI);
        const double fr = f * pd[q][6], fi = f * pd[q][5];
        bx += fr * dr - fi * di;
        by += fr * di + fi * dr;
    }
    coef[(int64_t)g * PG_TERMS + k] = make_double2(bx, by);
}

// panel_group_coeffs without the strengths: the group headers, and per
// panel p of the group and term k the factor geo[p * PG_TERMS + k] =
// R_g / (2 pi) (s_p + i c_p) (zh2^(k+1) - zh1^(k+1)) / (k + 1), so that
// b_k = sum_p gamma_p geo[p][k] (panel_group_apply).
__global__ __launch_bounds__(64) void panel_group_geom(
    int64_t M, const double *__restrict__ xs, const double *__restrict__ ys,
    const double *__restrict__ xe, const double *__restrict__ ye,
    const double *__restrict__ cosA, const double *__restrict__ sinA, PanelG *__restrict__ grp,
    double2 *__restrict__ geo)
{
    __shared__ double pd[PG_SIZE][6];  // xs, ys, xe, ye, cos, sin
    __shared__ double sc[3];
    const int g = blockIdx.x, k = threadIdx.x;
    const int64_t p0 = (int64_t)g * PG_SIZE;
    const int cnt = (int)min((int64_t)PG_SIZE, M - p0);
    if (k < cnt) {
        const int64_t p = p0 + k;
        pd[k][0] = xs[p];
        pd[k][1] = ys[p];
        pd[k][2] = xe[p];
        pd[k][3] = ye[p];
        pd[k][4] = cosA[p];
        pd[k][5] = sinA[p];
    }
    __syncthreads();
    if (k == 0) {
        double cx = 0.0, cy = 0.0;
        for (int q = 0; q < cnt; ++q) {
            cx += 0.5 * (pd[q][0] + pd[q][2]);
            cy += 0.5 * (pd[q][1] + pd[q][3]);
        }
        cx /= cnt;
        cy /= cnt;
        double r2 = 0.0;
        for (int q = 0; q < cnt; ++q) {
            const double a1 = pd[q][0] - cx, b1 = pd[q][1] - cy;
            const double a2 = pd[q][2] - cx, b2 = pd[q][3] - cy;
            r2 = fmax(r2, fmax(a1 * a1 + b1 * b1, a2 * a2 + b2 * b2));
        }
        const double R = sqrt(r2);
        sc[0] = cx;
        sc[1] = cy;
        sc[2] = R;
        PanelG pg;
        pg.zx = cx;
        pg.zy = cy;
        pg.far2 = 9.0 * r2;
        pg.R = R;
        pg.inv_R = 1.0 / R;
        pg.first = (int32_t)p0;
        pg.count = cnt;
        pg.pad0 = pg.pad1 = 0;
        grp[g] = pg;
    }
    __syncthreads();
    if (k >= PG_TERMS) return;
    const double cx = sc[0], cy = sc[1], R = sc[2], iR = 1.0 / R;
    for (int q = 0; q < cnt; ++q) {
        const double u1 = (pd[q][0] - cx) * iR, v1 = (pd[q][1] - cy) * iR;
        const double u2 = (pd[q][2] - cx) * iR, v2 = (pd[q][3] - cy) * iR;
        double p1r = u1, p1i = v1, p2r = u2, p2i = v2;
        for (int e = 0; e < k; ++e) {
            const double t1 = p1r * u1 - p1i * v1;
            p1i = p1r * v1 + p1i * u1;
            p1r = t1;
            const double t2 = p2r * u2 - p2i * v2;
            p2i = p2r * v2 + p2i * u2;
            p2r = t2;
        }
        const double dr = (p2r - p1r) / (k + 1), di = (p2i - p1i) / (k + 1);
        const double f = R / (2.0 * M_PI);
        const double fr = f * pd[q][5], fi = f * pd[q][4];
        geo[(p0 + q) * PG_TERMS + k] = make_double2(fr * dr - fi * di, fr * di + fi * dr);
    }
}

// b_k of every group from the strengths: one thread per (group, term),
// the group's panels in index order.
__global__ void panel_group_apply(int64_t M, int64_t NG, const double *__restrict__ gamma,
                                  const double2 *__restrict__ geo, double2 *__restrict__ coef)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NG * PG_TERMS) return;
    const int64_t g = e / PG_TERMS, k = e - g * PG_TERMS;
    const int64_t p0 = g * PG_SIZE, p1 = min(M, p0 + PG_SIZE);
    double bx = 0.0, by = 0.0;
    for (int64_t p = p0; p < p1; ++p) {
        const double2 f = geo[p * PG_TERMS + k];
        bx += gamma[p] * f.x;
        by += gamma[p] * f.y;
    }
    coef[e] = make_double2(bx, by);
}

int panel_series_geom_device(Lane *L, int64_t m, const double *xs, const double *ys,
                             const double *xe, const double *ye, const double *cosA,
                             const double *sinA, PanelSeriesGeom *out)
{
    if (m < 4 * PG_SIZE) {  // the per-panel path serves such bodies
        *out = PanelSeriesGeom{};
        return PHF_OK;
    }
    const int64_t ng = (m + PG_SIZE - 1) / PG_SIZE;
    PanelG *grp = out->m == m ? static_cast<PanelG *>(out->grp) : nullptr;
    double2 *geo = out->m == m ? out->geo : nullptr;
    if (!grp || !geo) {
        PHF_ALLOC(grp, L->arena, PanelG, (size_t)ng);
        PHF_ALLOC(geo, L->arena, double2, (size_t)m * PG_TERMS);
    }
    panel_group_geom<<<(unsigned)ng, 64, 0, L->stream>>>(m, xs, ys, xe, ye, cosA, sinA, grp, geo);
    PHF_HIP(hipGetLastError());
    out->grp = grp;
    out->geo = geo;
    out->m = m;
    return PHF_OK;
}

__global__ void target_keys16(const double *__restrict__ x, const double *__restrict__ y,
                              int64_t n, const double4 *__restrict__ bb,
                              uint32_t *__restrict__ keys, int32_t *__restrict__ vals)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 b = *bb;
    const double ex = fmax(b.y - b.x, 1e-300), ey = fmax(b.w - b.z, 1e-300);
    double fx = floor((x[i] - b.x) / ex * 256.0), fy = floor((y[i] - b.z) / ey * 256.0);
    fx = fmin(fmax(fx, 0.0), 255.0);
    fy = fmin(fmax(fy, 0.0), 255.0);
    keys[i] = morton2((uint32_t)fx, (uint32_t)fy);
    vals[i] = (int32_t)i;
}

template <bool F32>
__global__ __launch_bounds__(PI_T) void panel_induced_grouped(
    int64_t M, int64_t NG, const PanelK *__restrict__ pk, const PanelG *__restrict__ grp,
    const double2 *__restrict__ coef, int64_t ne, const int32_t *__restrict__ order,
    int ostride, const double *__restrict__ xE, const double *__restrict__ yE,
    double *__restrict__ vx, double *__restrict__ vy)
{
    constexpr int NT = F32 ? 16 : PG_TERMS;
    // series coefficients in the precision of the sum (float32: converted
    // once per chunk, not per term and target)
    using CT = typename std::conditional<F32, float2, double2>::type;
    __shared__ PanelG sg[PG_CHUNK];
    __shared__ CT sc[PG_CHUNK * NT];
    // PI_LPT lanes per target, each taking every PI_LPT-th group of a round:
    // the groups' chains are short and serial, so more of them in flight
    const int tid = threadIdx.x, sub = tid & (PI_LPT - 1);
    const int64_t k = (int64_t)blockIdx.x * (PI_T / PI_LPT) + tid / PI_LPT;
    const int64_t i = k < ne ? order[k * ostride] : 0;
    const double x = k < ne ? xE[i] : 0.0, y = k < ne ? yE[i] : 0.0;
    double sx = 0.0, sy = 0.0;
    for (int64_t g0 = 0; g0 < NG; g0 += PG_CHUNK) {
        const int ng = (int)min((int64_t)PG_CHUNK, NG - g0);
        __syncthreads();
        if (tid < ng) sg[tid] = grp[g0 + tid];
        for (int e = tid; e < ng * NT; e += PI_T) {
            const int gg = e / NT, kk = e - gg * NT;
            const double2 c = coef[(g0 + gg) * PG_TERMS + kk];
            if constexpr (F32)
                sc[e] = make_float2((float)c.x, (float)c.y);
            else
                sc[e] = c;
        }
        __syncthreads();
        for (int gg = sub; gg < ng; gg += PI_LPT) {
            const PanelG G = sg[gg];
            const double dx = x - G.zx, dy = y - G.zy;
            if (dx * dx + dy * dy > G.far2) {
                const CT *c = sc + gg * NT;
                double wr, wi;
                if constexpr (F32) {
                    // t = R / (z - z_g) in float32, Horner in four FMAs a term
                    const float fx = (float)dx, fy = (float)dy;
                    const float sc_t = (float)G.R * rcp_f32(__builtin_fmaf(fx, fx, fy * fy));
                    const float fr = fx * sc_t, fi = -fy * sc_t;
                    float ar = c[NT - 1].x, ai = c[NT - 1].y;
#pragma unroll
                    for (int q = NT - 2; q >= 0; --q) {
                        const float nr = __builtin_fmaf(ar, fr, __builtin_fmaf(-ai, fi, c[q].x));
                        ai = __builtin_fmaf(ar, fi, __builtin_fmaf(ai, fr, c[q].y));
                        ar = nr;
                    }
                    wr = (double)__builtin_fmaf(ar, fr, -ai * fi);
                    wi = (double)__builtin_fmaf(ar, fi, ai * fr);
                } else {
                    // t = R / (z - z_g)
                    const double d2 = dx * dx + dy * dy;
                    const double tr = G.R * dx / d2, ti = -G.R * dy / d2;
                    double ar = c[NT - 1].x, ai = c[NT - 1].y;
                    for (int q = NT - 2; q >= 0; --q) {
                        const double nr = ar * tr - ai * ti + c[q].x;
                        ai = ar * ti + ai * tr + c[q].y;
                        ar = nr;
                    }
                    wr = ar * tr - ai * ti;
                    wi = ar * ti + ai * tr;
                }
                // conj(v) = w / R
                sx += wr * G.inv_R;
                sy -= wi * G.inv_R;
            } else {
                // a near group's panels come from memory (L1/L2): only the
                // few targets beside the body read them
                for (int q = G.first; q < G.first + G.count; ++q) {
                    const PanelK p = pk[q];
                    const double ex = x - p.xs, ey = y - p.ys;
                    const double xP = p.c * ex + p.s * ey;
                    const double yP = -p.s * ex + p.c * ey;
                    const double xm = xP - p.x2;
                    const double cross = yP * p.x2, dot = xP * xm + yP * yP;
                    const double den = xm * xm + yP * yP;
                    const double num = p.x2 * (2.0 * xP - p.x2);
                    double th, lg;
                    if (F32) {
                        th = (double)atan2f((float)cross, (float)dot);
                        lg = (double)log1pf((float)num / (float)den);
                    } else {
                        th = atan2(cross, dot);
                        lg = log1p(num / den);
                    }
                    const double vxP = p.g2 * th;
                    const double vyP = -p.g4 * lg;
                    sx += p.c * vxP - p.s * vyP;
                    sy += p.s * vxP + p.c * vyP;
                }
            }
        }
    }
    // a target's lanes add their sums (a fixed butterfly)
#pragma unroll
    for (int o = 1; o < PI_LPT; o <<= 1) {
        sx += __shfl_xor(sx, o);
        sy += __shfl_xor(sy, o);
    }
    if (k < ne && sub == 0) {
        vx[i] = sx;
        vy[i] = sy;
    }
}

int panel_induced_device(Lane *L, int64_t m, const double *gamma, int64_t ne, const double *xE,
                         const double *yE, const double *xs, const double *ys, const double *xe,
                         const double *ye, const double *cosA, const double *sinA, double *vx,
                         double *vy, int prec, const PanelSeriesGeom *sg, const int32_t *order,
                         int order_stride)
{
    if (ne <= 0) return PHF_OK;
    PanelK *pk;
    PHF_ALLOC(pk, L->arena, PanelK, (size_t)std::max<int64_t>(m, 1));
    if (m > 0)
        panel_consts<<<div_up(m, 256), 256, 0, L->stream>>>(m, gamma, xs, ys, xe, ye, cosA, sinA,
                                                            pk);
    if (m < 4 * PG_SIZE || ne < 4096) {
        // few panels or targets: every pair by the per-panel formula
        if (prec == PHF_PREC_F32C)
            panel_induced_kernel<true><<<div_up(ne, PI_T), PI_T, 0, L->stream>>>(m, pk, ne, xE,
                                                                                 yE, vx, vy);
        else
            panel_induced_kernel<false><<<div_up(ne, PI_T), PI_T, 0, L->stream>>>(m, pk, ne, xE,
                                                                                  yE, vx, vy);
        PHF_HIP(hipGetLastError());
        return PHF_OK;
    }
    // groups and their series
    const int64_t ng = (m + PG_SIZE - 1) / PG_SIZE;
    PanelG *grp;
    double2 *coef;
    PHF_ALLOC(coef, L->arena, double2, (size_t)ng * PG_TERMS);
    if (sg && sg->grp && sg->m == m) {
        // fixed body: strengths times the precomputed factors
        grp = static_cast<PanelG *>(sg->grp);
        panel_group_apply<<<div_up(ng * PG_TERMS, 256), 256, 0, L->stream>>>(m, ng, gamma,
                                                                             sg->geo, coef);
    } else {
        PHF_ALLOC(grp, L->arena, PanelG, (size_t)ng);
        panel_group_coeffs<<<(unsigned)ng, 64, 0, L->stream>>>(m, gamma, xs, ys, xe, ye, cosA,
                                                               sinA, grp, coef);
    }
    // targets in Morton order: the caller's, or sorted here
    const int32_t *vs = order;
    int vstride = order_stride;
    if (!vs) {
        double4 *bb;
        PHF_ALLOC(bb, L->arena, double4, 1);
        PHF_TRY(device_bbox(L, xE, yE, ne, nullptr, nullptr, 0, bb));
        uint32_t *k0, *k1, *ks;
        int32_t *v0, *v1, *sv;
        PHF_ALLOC(k0, L->arena, uint32_t, (size_t)ne);
        PHF_ALLOC(k1, L->arena, uint32_t, (size_t)ne);
        PHF_ALLOC(v0, L->arena, int32_t, (size_t)ne);
        PHF_ALLOC(v1, L->arena, int32_t, (size_t)ne);
        target_keys16<<<div_up(ne, 256), 256, 0, L->stream>>>(xE, yE, ne, bb, k0, v0);
        PHF_HIP(hipGetLastError());
        PHF_TRY(device_radix_sort(L, k0, v0, k1, v1, ne, 16, &ks, &sv));
        vs = sv;
        vstride = 1;
    }
    if (prec == PHF_PREC_F32C)
        panel_induced_grouped<true><<<div_up(ne, PI_T / PI_LPT), PI_T, 0, L->stream>>>(
            m, ng, pk, grp, coef, ne, vs, vstride, xE, yE, vx, vy);
    else
        panel_induced_grouped<false><<<div_up(ne, PI_T / PI_LPT), PI_T, 0, L->stream>>>(
            m, ng, pk, grp, coef, ne, vs, vstride, xE, yE, vx, vy);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

int panel_influence_device(Lane *L, int64_t m, const double *xCP, const double *yCP,
                           const double *xs, const double *ys, const double *xe, const double *ye,
                           const double *cosA, const double *sinA, double *A)
{
    if (m <= 0) return PHF_OK;
    panel_influence_kernel<<<div_up(m * m, 256), 256, 0, L->stream>>>(m, xCP, yCP, xs, ys, xe, ye,
                                                                     cosA, sinA, A);
    PHF_HIP(hipGetLastError());
    return PHF_OK;
}

}  // namespace phf

// ------------------------------------------------------------ C entry ----
using namespace phf;

extern "C" {

int phf_panel_influence(const double *xCP, const double *yCP, const double *xs,
                        const double *ys, const double *xe, const double *ye,
                        const double *cosA, const double *sinA, int64_t m, double *A)
{
    clear_error();
    PHF_REQUIRE(m >= 0 && m <= 65536, "panel count out of range (%lld)", (long long)m);
    const double *in[8] = {xCP, yCP, xs, ys, xe, ye, cosA, sinA};
    const char *nm[8] = {"xCP", "yCP", "xPanelStart", "yPanelStart", "xPanelEnd", "yPanelEnd",
                         "cosAlpha", "sinAlpha"};
    for (int k = 0; k < 8; ++k) PHF_TRY(check_host_array(in[k], m, nm[k]));
    PHF_REQUIRE(m == 0 || A, "A is NULL");
    if (m == 0) return PHF_OK;
    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *L = g.lane;
    PHF_TRY(reset_arena(L));
    double *d[8], *dA;
    for (int k = 0; k < 8; ++k) PHF_TRY(upload(L, in[k], m, &d[k]));
    PHF_ALLOC(dA, L->arena, double, (size_t)(m * m));
    StageTimer tm(L, timing_enabled());
    bool used[PHF_STAGE_COUNT] = {};
    used[PHF_STAGE_PANEL] = true;
    tm.begin(PHF_STAGE_PANEL);
    panel_influence_kernel<<<div_up(m * m, 256), 256, 0, L->stream>>>(m, d[0], d[1], d[2], d[3],
                                                                     d[4], d[5], d[6], d[7], dA);
    PHF_HIP(hipGetLastError());
    tm.end(PHF_STAGE_PANEL);
    PHF_TRY(download(L, dA, m * m, A));
    PHF_HIP(hipStreamSynchronize(L->stream));
    record_stage_times(L, used);
    return PHF_OK;
}

int phf_panel_induced_velocity(const double *gamma, int64_t m, const double *xE,
                               const double *yE, int64_t ne, const double *xs,
                               const double *ys, const double *xe, const double *ye,
                               const double *cosA, const double *sinA, double *vx, double *vy)
{
    clear_error();
    PHF_REQUIRE(m >= 0 && ne >= 0, "negative sizes");
    const double *in[7] = {gamma, xs, ys, xe, ye, cosA, sinA};
    const char *nm[7] = {"gamma", "xPanelStart", "yPanelStart", "xPanelEnd", "yPanelEnd",
                         "cosAlpha", "sinAlpha"};
    for (int k = 0; k < 7; ++k) PHF_TRY(check_host_array(in[k], m, nm[k]));
    PHF_TRY(check_host_array(xE, ne, "xEval"));
    PHF_TRY(check_host_array(yE, ne, "yEval"));
    PHF_REQUIRE(ne == 0 || (vx && vy), "output arrays are NULL");
    if (ne == 0) return PHF_OK;
    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *L = g.lane;
    PHF_TRY(reset_arena(L));
    double *d[7], *dxE, *dyE, *dvx, *dvy;
    for (int k = 0; k < 7; ++k) PHF_TRY(upload(L, in[k], m, &d[k]));
    PHF_TRY(upload(L, xE, ne, &dxE));
    PHF_TRY(upload(L, yE, ne, &dyE));
    PHF_ALLOC(dvx, L->arena, double, (size_t)ne);
    PHF_ALLOC(dvy, L->arena, double, (size_t)ne);
    PHF_TRY(panel_induced_device(L, m, d[0], ne, dxE, dyE, d[1], d[2], d[3], d[4], d[5], d[6], dvx,
                                 dvy, PHF_PREC_F64));
    PHF_TRY(download(L, dvx, ne, vx));
    PHF_TRY(download(L, dvy, ne, vy));
    PHF_HIP(hipStreamSynchronize(L->stream));
    return PHF_OK;
}

int phf_dense_solve(const double *B, int64_t rows, int64_t n, const double *c, double *x,
                    phf_solve_info *info)
{
    clear_error();
    PHF_REQUIRE(n >= 1 && n <= 4096 && rows >= n && rows <= 65536,
                "matrix shape %lld x %lld unsupported (1 <= n <= 4096, n <= rows)",
                (long long)rows, (long long)n);
    PHF_TRY(check_host_array(B, rows * n, "A"));
    PHF_TRY(check_host_array(c, rows, "b"));
    PHF_REQUIRE(x != nullptr, "x is NULL");
    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *L = g.lane;
    PHF_TRY(reset_arena(L));
    auto t0 = std::chrono::steady_clock::now();
    double *dB, *dc, *dx;
    PHF_TRY(upload(L, B, rows * n, &dB));
    PHF_TRY(upload(L, c, rows, &dc));
    PHF_ALLOC(dx, L->arena, double, (size_t)n);
    SolveStats st{};
    int rc = dense_solve_device(L, dB, rows, n, dc, dx, &st);
    if (rc == PHF_OK) {
        PHF_TRY(download(L, dx, n, x));
        PHF_HIP(hipStreamSynchronize(L->stream));
    }
    auto t1 = std::chrono::steady_clock::now();
    if (info) {
        info->iterations = st.iterations;
        info->converged = st.converged ? 1 : 0;
        info->lu_ms = st.lu_ms;
        info->refine_ms = st.refine_ms;
        info->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    return rc;
}

int phf_lu_factor_f32(const float *A, int64_t n, float *LU, int32_t *ipiv, int32_t *info)
{
    clear_error();
    PHF_REQUIRE(n >= 1 && n <= 4096, "n out of range (%lld)", (long long)n);
    PHF_REQUIRE(A && LU && ipiv, "NULL argument");
    for (int64_t i = 0; i < n * n; ++i)
        PHF_REQUIRE(std::isfinite(A[i]), "A[%lld] is not finite", (long long)i);
    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *L = g.lane;
    PHF_TRY(reset_arena(L));
    float *dA;
    int32_t *dp, *di;
    PHF_ALLOC(dA, L->arena, float, (size_t)(n * n));
    PHF_ALLOC(dp, L->arena, int32_t, (size_t)n);
    PHF_ALLOC(di, L->arena, int32_t, 1);
    PHF_HIP(hipMemcpyAsync(dA, A, sizeof(float) * n * n, hipMemcpyHostToDevice, L->stream));
    PHF_TRY(lu_factor_f32(L, dA, (int)n, dp, di));
    PHF_HIP(hipMemcpyAsync(LU, dA, sizeof(float) * n * n, hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipMemcpyAsync(ipiv, dp, sizeof(int32_t) * n, hipMemcpyDeviceToHost, L->stream));
    int32_t hinfo = 0;
    PHF_HIP(hipMemcpyAsync(&hinfo, di, sizeof(int32_t), hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    if (info) *info = hinfo;
    return PHF_OK;
}

int phf_chol_factor_f32(const float *A, int64_t n, float *LU, int32_t *info)
{
    clear_error();
    PHF_REQUIRE(n >= 1 && n <= 1024, "n out of range (%lld)", (long long)n);
    PHF_REQUIRE(A && LU, "NULL argument");
    for (int64_t i = 0; i < n * n; ++i)
        PHF_REQUIRE(std::isfinite(A[i]), "A[%lld] is not finite", (long long)i);
    LaneGuard g;
    PHF_TRY(g.rc);
    Lane *L = g.lane;
    PHF_TRY(reset_arena(L));
    float *dA, *dg;
    int32_t *dp, *di;
    PHF_ALLOC(dA, L->arena, float, (size_t)(n * n));
    PHF_ALLOC(dg, L->arena, float, (size_t)n);
    PHF_ALLOC(dp, L->arena, int32_t, (size_t)n);
    PHF_ALLOC(di, L->arena, int32_t, 1);
    PHF_HIP(hipMemcpyAsync(dA, A, sizeof(float) * n * n, hipMemcpyHostToDevice, L->stream));
    PHF_TRY(chol_factor_f32(L, dA, (int)n, dp, di, dg));
    PHF_HIP(hipMemcpyAsync(LU, dA, sizeof(float) * n * n, hipMemcpyDeviceToHost, L->stream));
    int32_t hinfo = 0;
    PHF_HIP(hipMemcpyAsync(&hinfo, di, sizeof(int32_t), hipMemcpyDeviceToHost, L->stream));
    PHF_HIP(hipStreamSynchronize(L->stream));
    if (info) *info = hinfo;
    return PHF_OK;
}

}  // extern "C"
