// als.hip -- CP-ALS sweep kernels and driver for gfx950.
//
// One mode update (reference src/cp_als.cu:46-63, 66-82, 85-101; src/cp4_als.cu:67-151):
//   V  = hadamard of the other modes' Grams           (FP64, R x R)
//   M  = MTTKRP_n(X, other factors)                    (fused kernel, split-K FP32 partials)
//   U_n = M V^{-1}                                      (FP64, written FP32 column-major
//                                                        plus the panel layout the next
//                                                        MTTKRP reads)
//   G_n = U_n^T U_n                                     (FP64, fixed-order reduction)
// The reference does the same with a KR batched GEMM, an SGEMM on an unfolded copy of X,
// FP32 LU (getrf/getrs) with a cudaMalloc per call, and a transpose (Sgeam).
//
// Kernels per mode: k_vinv (one workgroup, symmetric sweep-operator inverse, Jacobi
// pseudo-inverse fallback), the fused MTTKRP (or the shared pass, order 3 BF16/FP8), then
// k_solve (one workgroup per 16 rows: FP64 sum of the split-K partials, U = M V^{-1}, panel,
// fit and Gram partials) and k_gram_psum (Gram of the new factor; at R <= 16 summed by the
// next mode's k_vinv instead); after the last mode k_gram_psum_fit (its Gram and the
// sweep's fit in one kernel).
// A sweep is captured once into a hipGraph and replayed.
#include "mttkrp.hpp"
#include "linalg.hpp"
#include "vinv.hpp"
#include <memory>
#include <mutex>
#include "mlp.hpp"
#include <math.h>
#include <string.h>
#include <vector>

namespace tdx {

static constexpr int kSolveRows = 16;  // rows per k_solve workgroup
static constexpr int kGramRows = 64;   // rows per Gram-partial workgroup

// gram_part != nullptr: first G_gmode (Gsum) = the ordered sum of the nblk Gram partials of
// that mode (what k_gram_psum computes, same order), so a small-rank sweep, whose inverses
// run in line, needs no separate Gram kernel between a solve and the next mode's inverse.
template <int NS>
__global__ __launch_bounds__(256) void k_vinv(const double *G0, const double *G1, const double *G2, const double *G3,
                                              int order, int mode, int R, double *__restrict__ Vinv,
                                              int *__restrict__ status, double *__restrict__ fb_ws,
                                              const double *__restrict__ gram_part, int nblk, double *Gsum) {
    const double *G[4] = {G0, G1, G2, G3};
    if (gram_part) {
        for (int e = threadIdx.x; e < R * R; e += blockDim.x) Gsum[e] = ordered_sum(gram_part + e, (long)R * R, nblk);
        __syncthreads();
    }
    vinv_block<NS>(G, order, mode, R, Vinv, status, fb_ws);
}

// M = sum of the split-K partials of the MTTKRP, FP64, in split order (deterministic);
// one thread per element of the rows_pad x Rp block, 8 independent loads in flight
__global__ __launch_bounds__(256) void k_split_sum(const float *__restrict__ partial, int nsplit, long nel,
                                                   double *__restrict__ msum) {
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= nel) return;
    double s = 0.0;
    int sp = 0;
    for (; sp + 8 <= nsplit; sp += 8) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = partial[(long)(sp + u) * nel + e];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += (double)v[u];
    }
    for (; sp < nsplit; ++sp) s += (double)partial[(long)sp * nel + e];
    msum[e] = s;
}

// U_n rows = M V^{-1} for kSolveRows rows per workgroup (256 threads), M summed here from
// the MTTKRP's split partials in FP64 in split order (the order of k_split_sum).  Also
// writes the panel layout (optionally row-scaled for FP8 storage), the per-block inner
// product <M, U> (fit) and the per-block Gram partial U_blk^T U_blk.  Register blocks:
// 1 x 4 outputs of U per thread (V^{-1} is symmetric, so its rows are read contiguously),
// 4 x 4 of the Gram partial.
__global__ __launch_bounds__(256) void k_solve(const float *__restrict__ partial, int nsplit, long rows_pad, int Rp,
                                               long In, int R, const double *__restrict__ Vinv,
                                               float *__restrict__ U, float *__restrict__ panel,
                                               const float *__restrict__ panel_scale, double *__restrict__ ip_part,
                                               double *__restrict__ gram_part, int *__restrict__ skip_reset,
                                               int quant) {
    __shared__ double Vs[kMaxR * kMaxR];
    __shared__ double Ms[kSolveRows * kMaxR];  // M rows, then the U rows (as rounded FP32 values)
    __shared__ double sh[256];
    const int tid = threadIdx.x;
    const long row0 = (long)blockIdx.x * kSolveRows;
    const int RR = R * R;
    // V^{-1} into registers (all of a thread's loads in flight at once), the split sums, then
    // V^{-1} into LDS: the two read phases overlap instead of chaining their latencies
    constexpr int kVPer = kMaxR * kMaxR / 256;
    double vreg[kVPer];
#pragma unroll
    for (int k = 0; k < kVPer; ++k) {
        const int e = tid + 256 * k;
        vreg[k] = e < RR ? Vinv[e] : 0.0;
    }
    // thread (row r, columns b .. b + 3): one 16-byte load per split (Rp % 16 == 0, so
    // kSolveRows Rp / 4 <= 256 threads cover the block), 16 splits in flight
    const int q4 = Rp / 4;
    if (tid < kSolveRows * q4) {
        const int r = tid / q4, b = 4 * (tid % q4);
        const long i = row0 + r;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        if (i < In) {
            const float4 *pp = reinterpret_cast<const float4 *>(partial + i * Rp + b);
            const long ps4 = rows_pad * Rp / 4;
            for (int sp = 0; sp < nsplit; sp += 16) {
                float4 v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int sq = sp + u < nsplit ? sp + u : nsplit - 1;  // valid; not summed
                    v[u] = pp[(long)sq * ps4];
                }
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (sp + u < nsplit) {  // split order, as k_split_sum
                        acc[0] += (double)v[u].x;
                        acc[1] += (double)v[u].y;
                        acc[2] += (double)v[u].z;
                        acc[3] += (double)v[u].w;
                    }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) Ms[r * kMaxR + b + q] = acc[q];
    }
#pragma unroll
    for (int k = 0; k < kVPer; ++k) {
        const int e = tid + 256 * k;
        if (e < RR) Vs[e] = vreg[k];
    }
    __syncthreads();
    // U = M V^{-1}: thread (row r, columns b0 .. b0 + 3)
    const int r = tid >> 4, b0 = 4 * (tid & 15);
    const long i = row0 + r;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (b0 < R) {
        for (int a = 0; a < R; ++a) {
            const double m = Ms[r * kMaxR + a];
            const double *vr = Vs + a * R + b0;  // (V^{-1})[a][b0..b0+3] (past R: unused)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[q] = fma(m, vr[q], acc[q]);
        }
    }
    // quant: the factor becomes the BF16 value of its panel entry (u sc rounded to BF16,
    // divided by sc), so that the factor, its Gram and the panel the shared pass reads as
    // BF16 MFMA operands describe the same iterate
    const float sc = (panel_scale && i < In) ? panel_scale[i] : 1.0f;
    float uf[4], pf[4];
    double ip = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const bool in = i < In && b0 + q < R;
        uf[q] = in ? (float)acc[q] : 0.0f;
        pf[q] = uf[q] * sc;
        if (quant) {
            pf[q] = (float)(__bf16)pf[q];
            uf[q] = pf[q] / sc;
        }
        if (in) ip += Ms[r * kMaxR + b0 + q] * (double)uf[q];
    }
    __syncthreads();  // every thread is done reading M
    if (b0 < Rp) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int b = b0 + q;
            if (b >= Rp) break;
            Ms[r * kMaxR + b] = (double)uf[q];
            if (i < In && b < R) U[i + (long)b * In] = uf[q];
            panel[((i >> 3) * Rp + b) * 8 + (i & 7)] = pf[q];
        }
    }
    ip = block_sum(ip, sh);  // (its barriers also publish the U rows)
    if (tid == 0) ip_part[blockIdx.x] = ip;
    // the mode-0 MTTKRP that may have been skipped has already run (stream order)
    if (skip_reset && blockIdx.x == 0 && tid == 0) *skip_reset = 0;
    // Gram partial: thread (rows a0 .. a0 + 3, columns c0 .. c0 + 3)
    const int a0 = 4 * (tid >> 4), c0 = 4 * (tid & 15);
    if (a0 < R && c0 < R) {
        double g[4][4];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) g[p][q] = 0.0;
#pragma unroll 4
        for (int rr = 0; rr < kSolveRows; ++rr) {
            double ua[4], ub[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                ua[q] = Ms[rr * kMaxR + a0 + q];
                ub[q] = Ms[rr * kMaxR + c0 + q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) g[p][q] = fma(ua[p], ub[q], g[p][q]);
        }
        double *gp = gram_part + (long)blockIdx.x * RR;
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int p = 0; p < 4; ++p)
                if (a0 + p < R && c0 + q < R) gp[(a0 + p) + (c0 + q) * R] = g[p][q];
    }
}


// G = sum of the block Gram partials, in block order; one thread per entry
__global__ __launch_bounds__(256) void k_gram_psum(const double *__restrict__ gram_part, int nblk, int RR,
                                                   double *__restrict__ Gout) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= RR) return;
    Gout[e] = ordered_sum(gram_part + e, RR, nblk);
}

// The fast-fit relative error of Tensorlab cpd_als.m:281-283,
//   ||X - Xhat||^2 = ||X||^2 - 2 <M, U_n> + sum(hadamard of all Grams),
// appended to fit_hist[*counter]; *counter is incremented.

// Gram partials U_blk^T U_blk (FP64) of kGramRows-row blocks of a column-major I x R factor
// The last mode's Gram (as k_gram_psum) fused with the sweep's fit: each block writes its
// Gram entries and the partial sum of the Hadamard product of all Grams over them; the last
// block to finish (atomic ticket, reset by that block) adds the partials in block order and
// <M, U> (ip_part, in block order) and appends the relative error to fit_hist.
__global__ __launch_bounds__(256) void k_gram_psum_fit(const double *__restrict__ gram_part, int nblk, int R,
                                                       int order, int mode, double *Gout, const double *G0,
                                                       const double *G1, const double *G2, const double *G3,
                                                       const double *__restrict__ ip_part, int nblk_ip,
                                                       const double *__restrict__ normx2, double *__restrict__ fit_hist,
                                                       int *__restrict__ counter, double *__restrict__ hpart,
                                                       int *__restrict__ done) {
    __shared__ double sh[256];
    __shared__ int last;
    const int RR = R * R, tid = threadIdx.x;
    const int e = blockIdx.x * 256 + tid;
    const double *G[4] = {G0, G1, G2, G3};
    double h = 0.0;
    if (e < RR) {
        // the other modes' Gram entries (straight-line loads, in flight with the partials'),
        // multiplied into the sum in mode order
        const double *Ga = (mode == 0) ? G[1] : G[0];
        const double *Gb = (mode <= 1) ? G[2] : G[1];
        const double *Gc = (order > 3) ? ((mode <= 2) ? G[3] : G[2]) : Ga;
        const double g0 = Ga[e], g1 = Gb[e], g2 = Gc[e];
        const double s = ordered_sum(gram_part + e, RR, nblk);
        Gout[e] = s;
        h = s * g0 * g1;
        if (order > 3) h *= g2;
    }
    h = block_sum(h, sh);
    if (tid == 0) {
        hpart[blockIdx.x] = h;
        __threadfence();
        last = (atomicAdd(done, 1) == (int)gridDim.x - 1);
    }
    __syncthreads();
    if (!last) return;
    if (tid == 0) {
        __threadfence();
        const double hx = ordered_sum(hpart, 1, (int)gridDim.x), ip = ordered_sum(ip_part, 1, nblk_ip);
        const double nx2 = *normx2;
        const double r2 = nx2 - 2.0 * ip + hx;
        const int c = *counter;
        fit_hist[c] = sqrt(fabs(r2) / nx2);
        *counter = c + 1;
        *done = 0;
    }
}

__global__ __launch_bounds__(256) void k_gram_rows(const float *__restrict__ U, long I, int R,
                                                   double *__restrict__ gram_part) {
    __shared__ double Us[kGramRows * kMaxR];
    const long row0 = (long)blockIdx.x * kGramRows;
    for (int e = threadIdx.x; e < kGramRows * R; e += 256) {
        const int r = e % kGramRows, b = e / kGramRows;
        const long i = row0 + r;
        Us[r * kMaxR + b] = (i < I) ? (double)U[i + (long)b * I] : 0.0;
    }
    __syncthreads();
    const int RR = R * R;
    double *gp = gram_part + (long)blockIdx.x * RR;
    for (int e = threadIdx.x; e < RR; e += 256) {
        const int a = e % R, b = e / R;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
        for (int r = 0; r < kGramRows; r += 4) {
#pragma unroll
            for (int q = 0; q < 4; ++q) s[q] += Us[(r + q) * kMaxR + a] * Us[(r + q) * kMaxR + b];
        }
        gp[e] = (s[0] + s[1]) + (s[2] + s[3]);
    }
}

// G = U^T U (FP64) of a column-major I x R factor: row-block partials, then their sum
static int gram_of(const float *U, long I, int R, double *gram_part, double *G, hipStream_t s) {
    const int nb = (int)ceil_div(I, kGramRows);
    k_gram_rows<<<nb, 256, 0, s>>>(U, I, R, gram_part);
    k_gram_psum<<<(int)ceil_div((long)R * R, 256L), 256, 0, s>>>(gram_part, nb, R * R, G);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

// ||X||^2 in FP64 with Kahan compensation per thread; per-block partials then one block
template <typename T>
__global__ __launch_bounds__(256) void k_norm2_partial(const T *__restrict__ X, long n, const float *__restrict__ scales,
                                                       long I, double *__restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0, c = 0.0;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        double x = (double)X[e];
        if (scales) x *= (double)scales[e % I];
        double y = x * x - c, t = s + y;
        c = (t - s) - y;
        s = t;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ||X||^2 in FP64 (per-thread Kahan over float4 groups, block partials) with 16-byte
// nontemporal loads, four in flight per thread; with Q the same pass writes the BF16 copy of X
// (round to nearest even, the bits tdx_quantize_bf16 writes).  X and Xq 16-byte aligned.
// A (FP8 storage): also the per-mode-1-slice absolute maxima (bit patterns, NaN excluded, as
// k_slice_amax) into amax[I]; the launch makes the grid stride (in elements) a multiple of I,
// with I % 4 == 0, so a thread's float4s all cover the same 4 slices and it issues one
// atomicMax per slice at the end.
template <bool Q, bool A = false>
__global__ __launch_bounds__(256) void k_norm2_stream(const float *__restrict__ X, long n, uint16_t *__restrict__ Xq,
                                                      double *__restrict__ part, uint32_t *__restrict__ amax = nullptr,
                                                      long I = 1) {
    __shared__ double sh[256];
    double s = 0.0, c = 0.0;
    uint32_t mx[4] = {0u, 0u, 0u, 0u};
    auto add = [&](double y) {
        y -= c;
        const double t = s + y;
        c = (t - s) - y;
        s = t;
    };
    auto take = [&](const float4 &v, long i4) {
        add((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z + (double)v.w * v.w);
        if constexpr (A) {
            const float w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t u = __float_as_uint(w[q]) & 0x7fffffffu;
                if (u > 0x7f800000u) u = 0u;
                mx[q] = u > mx[q] ? u : mx[q];
            }
        }
        if constexpr (Q)
            reinterpret_cast<uint2 *>(Xq)[i4] =
                uint2{(uint32_t)f32_to_bf16_bits(v.x) | ((uint32_t)f32_to_bf16_bits(v.y) << 16),
                      (uint32_t)f32_to_bf16_bits(v.z) | ((uint32_t)f32_to_bf16_bits(v.w) << 16)};
    };
    const float4 *X4 = reinterpret_cast<const float4 *>(X);
    const long n4 = n >> 2, stride = (long)gridDim.x * 256;
    long i = blockIdx.x * 256L + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = ld_stream(X4 + i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u) take(v[u], i + u * stride);
    }
    for (; i < n4; i += stride) take(ld_stream(X4 + i), i);
    if (blockIdx.x == 0)
        for (long e = 4 * n4 + threadIdx.x; e < n; e += 256) {  // (none when A: I % 4 == 0)
            add((double)X[e] * X[e]);
            if constexpr (Q) Xq[e] = f32_to_bf16_bits(X[e]);
        }
    if constexpr (A) {
        const long i0 = blockIdx.x * 256L + threadIdx.x;
        if (i0 < n4) {
            const long sl = (4 * i0) % I;
#pragma unroll
            for (int q = 0; q < 4; ++q) atomicMax(&amax[sl + q], mx[q]);
        }
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_sum_partials(const double *__restrict__ part, int n, double *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = s;
}

__global__ void k_set_int(int *p, int v) { *p = v; }
__global__ void k_fill_u32_dev(uint32_t *p, long n, uint32_t v) {
    for (long i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------------
// Extrapolation step (Tensorlab cpd_als.m:234-339 line search with the step length from
// the learned MLP, the cpd_lsb.m:30 rule, or a grid for training labels).  After the plain
// sweep U_start -> U, U_a = U_start + alpha (U - U_start) is evaluated exactly:
// <MTTKRP_0(U_a), U_a,0> and the Grams of U_a give ||X - [[U_a]]||^2 in FP64.  It replaces
// U only if it is better; the mode-0 MTTKRP of U_a then serves the next sweep's first update.

struct FacSet {
    float *p[4];
    long I[4];
    int order;
    int R;
};

struct LsState {
    int tries;
    int accepts;
    float last_alpha;
    float alpha_cur;
    int accept;
    int skip;
};

__device__ inline void fac_locate(const FacSet &f, long e, int &m, long &off) {
    m = 0;
    off = e;
    while (m < f.order - 1 && off >= f.I[m] * f.R) {
        off -= f.I[m] * f.R;
        ++m;
    }
}

__global__ void k_ls_init(LsState *st) {
    st->tries = 0;
    st->accepts = 0;
    st->last_alpha = 1.0f;
    st->alpha_cur = 1.0f;
    st->accept = 0;
    st->skip = 0;
}

__global__ void k_copy_set(FacSet src, FacSet dst) {
    long total = 0;
    for (int m = 0; m < src.order; ++m) total += src.I[m] * src.R;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        int m;
        long off;
        fac_locate(src, e, m, off);
        dst.p[m][off] = src.p[m][off];
    }
}

// features (see TDX_LS_FEATURES in tdx.h) and the step length; one workgroup
__global__ __launch_bounds__(256) void k_ls_features(FacSet U, FacSet U1, const double *__restrict__ fit_hist,
                                                     const int *__restrict__ counter, LsState *__restrict__ st,
                                                     int mode, float grid_alpha, const float *__restrict__ mlp,
                                                     int width, float *__restrict__ feat_hist,
                                                     float *__restrict__ alpha_hist, double *__restrict__ plain_hist) {
    __shared__ double sh[256];
    __shared__ float feat[TDX_LS_FEATURES];
    __shared__ float msh[256];
    const int tid = threadIdx.x;
    const int c = *counter;  // sweeps recorded so far; this sweep is c - 1
    const int k = c - 1;
    double chg[4] = {0, 0, 0, 0};
    for (int m = 0; m < U.order; ++m) {
        double d2 = 0.0, u2 = 0.0;
        const long n = U.I[m] * U.R;
        for (long e = tid; e < n; e += 256) {
            const double u = U.p[m][e], d = u - (double)U1.p[m][e];
            d2 += d * d;
            u2 += u * u;
        }
        d2 = block_sum(d2, sh);
        u2 = block_sum(u2, sh);
        chg[m] = sqrt(d2 / fmax(u2, 1e-300));
    }
    if (tid == 0) {
        const double r0 = fit_hist[k];
        feat[0] = (float)log10(fmax(r0, 1e-30));
        for (int j = 1; j <= 3; ++j) {
            float v = 0.0f;
            if (k - j >= 0) v = (float)log10(fmax(fit_hist[k - j], 1e-30) / fmax(fit_hist[k - j + 1], 1e-30));
            feat[j] = v;
        }
        for (int m = 0; m < 4; ++m) feat[4 + m] = (m < U.order) ? (float)log10(chg[m] + 1e-12) : 0.0f;
        feat[8] = (float)log10((double)k + 1.0);
        feat[9] = (U.order == 4) ? 1.0f : 0.0f;
        feat[10] = logf(fmaxf(st->last_alpha, 1.0f));
        feat[11] = (st->tries > 0) ? (float)st->accepts / (float)st->tries : 0.0f;
        if (plain_hist) plain_hist[k] = r0;
    }
    __syncthreads();
    float alpha;
    if (mode == 1) {
        MlpShape sh_{TDX_LS_FEATURES, width, 1};
        const float beta = mlp_forward_one(mlp, sh_, feat, msh);
        alpha = expf(fminf(fmaxf(beta, 0.0f), 3.4657359f));  // [1, 32]
    } else if (mode == 2) {
        alpha = (float)cbrt((double)k + 1.0);
    } else {
        alpha = grid_alpha;
    }
    if (tid == 0) {
        st->alpha_cur = alpha;
        if (alpha_hist) alpha_hist[k] = alpha;
    }
    if (feat_hist && tid < TDX_LS_FEATURES) feat_hist[(long)k * TDX_LS_FEATURES + tid] = feat[tid];
}

// U_a = U_start + alpha (U - U_start), written column-major and as panels (mode-0 panel
// scaled by the FP8 slice scales when given)
__global__ void k_extrapolate(FacSet U, FacSet U1, FacSet Ua, float *pa0, float *pa1, float *pa2, float *pa3, int Rp,
                              const LsState *__restrict__ st, const float *__restrict__ scale0) {
    float *pa[4] = {pa0, pa1, pa2, pa3};
    const float alpha = st->alpha_cur;
    long total = 0, prow[4];
    for (int m = 0; m < U.order; ++m) {
        prow[m] = (U.I[m] + 63) / 64 * 64;
        total += prow[m] * Rp;
    }
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        int m = 0;
        long off = e;
        while (m < U.order - 1 && off >= prow[m] * Rp) {
            off -= prow[m] * Rp;
            ++m;
        }
        const long i = ((off >> 3) / Rp) * 8 + (off & 7);
        const long r = (off >> 3) % Rp;
        float v = 0.0f;
        if (i < U.I[m] && r < U.R) {
            const long idx = i + r * U.I[m];
            const float u1 = U1.p[m][idx];
            v = fmaf(alpha, U.p[m][idx] - u1, u1);
            Ua.p[m][idx] = v;
            if (m == 0 && scale0) v *= scale0[i];
        }
        pa[m][off] = v;
    }
}

// Grams of a factor set, one workgroup per mode
// per-block <M_0(U_a), U_a,0> with M_0 the FP64 sum of the split partials (msum)
__global__ __launch_bounds__(256) void k_ls_ip(const double *__restrict__ msum, int Rp, long In, int R,
                                               const float *__restrict__ Ua0, double *__restrict__ ip_part) {
    __shared__ double sh[256];
    const long row0 = (long)blockIdx.x * kSolveRows;
    double ip = 0.0;
    for (int e = threadIdx.x; e < kSolveRows * Rp; e += 256) {
        const long i = row0 + e / Rp;
        const int b = e % Rp;
        if (i >= In || b >= R) continue;
        ip += msum[i * Rp + b] * (double)Ua0[i + (long)b * In];
    }
    ip = block_sum(ip, sh);
    if (threadIdx.x == 0) ip_part[blockIdx.x] = ip;
}

// exact relative error of U_a; accept (line_search 1, 2) or record (3)
__global__ __launch_bounds__(256) void k_ls_decide(const double *__restrict__ ip_part, int nblk, const double *Ga0,
                                                   const double *Ga1, const double *Ga2, const double *Ga3, int order,
                                                   int R, const double *__restrict__ normx2, double *fit_hist,
                                                   const int *__restrict__ counter, int mode, int g, int ngrid,
                                                   double *__restrict__ grid_hist, int *__restrict__ accepted_hist,
                                                   LsState *__restrict__ st) {
    __shared__ double sh[256];
    const double *Ga[4] = {Ga0, Ga1, Ga2, Ga3};
    double hx = 0.0;
    for (int e = threadIdx.x; e < R * R; e += 256) {
        double v = 1.0;
        for (int m = 0; m < order; ++m) v *= Ga[m][e];
        hx += v;
    }
    hx = block_sum(hx, sh);
    double ip = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) ip += ip_part[b];
    ip = block_sum(ip, sh);
    if (threadIdx.x != 0) return;
    const double nx2 = *normx2;
    const double rel_a = sqrt(fabs(nx2 - 2.0 * ip + hx) / nx2);
    const int k = *counter - 1;
    if (mode == 3) {
        if (grid_hist) grid_hist[(long)k * ngrid + g] = rel_a;
        st->accept = 0;
        st->skip = 0;
        return;
    }
    const int acc = (isfinite(rel_a) && rel_a < fit_hist[k]) ? 1 : 0;
    st->tries += 1;
    if (acc) {
        st->accepts += 1;
        st->last_alpha = st->alpha_cur;
        fit_hist[k] = rel_a;
    } else {
        st->last_alpha = 1.0f;
    }
    st->accept = acc;
    st->skip = acc;
    if (accepted_hist) accepted_hist[k] = acc;
}

// on acceptance: U <- U_a, panels <- panels_a, G <- G_a
__global__ void k_ls_commit(FacSet U, FacSet Ua, float *p0, float *p1, float *p2, float *p3, const float *pa0,
                            const float *pa1, const float *pa2, const float *pa3, int Rp, double *G0, double *G1,
                            double *G2, double *G3, const double *Ga0, const double *Ga1, const double *Ga2,
                            const double *Ga3, const LsState *__restrict__ st) {
    if (!st->accept) return;
    float *pd[4] = {p0, p1, p2, p3};
    const float *ps[4] = {pa0, pa1, pa2, pa3};
    double *Gd[4] = {G0, G1, G2, G3};
    const double *Gs[4] = {Ga0, Ga1, Ga2, Ga3};
    const long gstride = (long)gridDim.x * blockDim.x, g0 = blockIdx.x * (long)blockDim.x + threadIdx.x;
    for (int m = 0; m < U.order; ++m) {
        const long n = U.I[m] * U.R;
        for (long e = g0; e < n; e += gstride) U.p[m][e] = Ua.p[m][e];
        const long np = (U.I[m] + 63) / 64 * 64 * Rp;
        for (long e = g0; e < np; e += gstride) pd[m][e] = ps[m][e];
        for (long e = g0; e < (long)U.R * U.R; e += gstride) Gd[m][e] = Gs[m][e];
    }
}

// ---------------------------------------------------------------------------
// host driver

struct CpPlan {
    int order = 0;
    long dims[4] = {1, 1, 1, 1};
    long R = 0;
    int Rp = 0;
    float *U[4] = {nullptr, nullptr, nullptr, nullptr};
    // device allocations (owned)
    float *panel[4] = {nullptr, nullptr, nullptr, nullptr};
    double *G[4] = {nullptr, nullptr, nullptr, nullptr};
    double *Vinv[4] = {nullptr, nullptr, nullptr, nullptr};
    float *partial = nullptr;
    size_t partial_floats = 0;
    double *msum = nullptr;  // FP64 sum of the partials, rows_pad x Rp
    size_t msum_doubles = 0;
    double *ip_part = nullptr;
    double *gram_part = nullptr;
    double *fit_hist = nullptr;
    int *counter = nullptr;
    int *status = nullptr;
    double *normx2 = nullptr;
    double *scratch = nullptr;
    double *vinv_ws = nullptr;  // k_vinv fallback workspace, 2 R^2 doubles
    double *hpart = nullptr;    // k_gram_psum_fit: per-block partials (R^2 / 256 doubles)
    int *fit_done = nullptr;    // k_gram_psum_fit: ticket (0 between launches)
    bool shared_pass = true;    // modes 1 and 2 from one pass over X where eligible
    uint16_t *T = nullptr;      // shared-pass intermediate X x_0 A^T (BF16, I2 x I1 x Rp)
    size_t t_bytes = 0;
    int max_sweeps_hist = 0;
    hipStream_t side = nullptr;
    hipEvent_t ev_fork[4], ev_join[4];
    bool events = false;
    // extrapolation step
    int ls_mode = 0;
    const float *mlp = nullptr;
    int mlp_width = 64;
    int ngrid = 0;
    double grid[8];
    float *U1[4] = {nullptr, nullptr, nullptr, nullptr};
    float *Ua[4] = {nullptr, nullptr, nullptr, nullptr};
    float *panel_a[4] = {nullptr, nullptr, nullptr, nullptr};
    double *Ga[4] = {nullptr, nullptr, nullptr, nullptr};
    LsState *ls = nullptr;
    double *ip_part_a = nullptr;
    float *feat_hist = nullptr, *alpha_hist = nullptr;
    int *accepted_hist = nullptr;
    double *grid_hist = nullptr, *plain_hist = nullptr;
    std::vector<void *> owned;
    // capacities of the shape-dependent buffers (a later call of another shape reuses the plan
    // when they suffice)
    long cap_panel[4] = {0, 0, 0, 0};  // floats per mode panel (and per mode of U1 / Ua)
    long cap_R = 0;                    // R x R buffers hold cap_R^2 doubles
    long cap_nblk = 0;                 // solve blocks of the longest mode

    ~CpPlan() { release(); }
    FacSet set_of(float *const *p) const {
        FacSet f{};
        f.order = order;
        f.R = (int)R;
        for (int m = 0; m < 4; ++m) {
            f.p[m] = m < order ? p[m] : nullptr;
            f.I[m] = m < order ? dims[m] : 0;
        }
        return f;
    }
    void release() {
        for (void *p : owned)
            if (p) (void)hipFree(p);
        owned.clear();
        for (int m = 0; m < 4; ++m) {
            if (panel[m]) (void)hipFree(panel[m]);
            if (G[m]) (void)hipFree(G[m]);
            if (Vinv[m]) (void)hipFree(Vinv[m]);
            panel[m] = nullptr; G[m] = nullptr; Vinv[m] = nullptr;
        }
        void *ps[] = {partial, msum, ip_part, gram_part, fit_hist, counter, status, normx2, scratch, vinv_ws, T,
                      hpart, fit_done};
        for (void *p : ps)
            if (p) (void)hipFree(p);
        partial = nullptr; msum = nullptr; ip_part = nullptr; gram_part = nullptr; fit_hist = nullptr;
        counter = nullptr; status = nullptr; normx2 = nullptr; scratch = nullptr; vinv_ws = nullptr; T = nullptr;
        hpart = nullptr; fit_done = nullptr;
        t_bytes = 0;
        if (events) {
            for (int m = 0; m < 4; ++m) {
                (void)hipEventDestroy(ev_fork[m]);
                (void)hipEventDestroy(ev_join[m]);
            }
            events = false;
        }
        if (side) (void)hipStreamDestroy(side);
        side = nullptr;
    }
};

struct Storage {
    const void *X;
    int dtype;
    const float *scales;  // FP8 slice scales
    MttkrpPlan mp[4];
    DtreePlan dt;         // order 3, BF16/FP8: modes 1 and 2 from one shared pass over X
};

// sets the shape of a plan whose buffers are (or will be) large enough for it
static void plan_bind(CpPlan &P, int order, const long *dims, long R, float *const *factors) {
    P.order = order;
    for (int m = 0; m < 4; ++m) P.dims[m] = m < order ? dims[m] : 1;
    P.R = R;
    P.Rp = (int)(ceil_div(R, 16) * 16);
    for (int m = 0; m < 4; ++m) P.U[m] = m < order ? factors[m] : nullptr;
}

// whether a cached plan's buffers hold this shape
static bool plan_fits(const CpPlan &P, int order, const long *dims, long R, int max_hist) {
    if (P.order != order || R > P.cap_R || max_hist > P.max_sweeps_hist) return false;
    const int Rp = (int)(ceil_div(R, 16) * 16);
    long nblk_max = 0;
    for (int m = 0; m < order; ++m) {
        if (panel_floats(dims[m], Rp) > P.cap_panel[m] || dims[m] * R > P.cap_panel[m]) return false;
        nblk_max = std::max(nblk_max, ceil_div(dims[m], kSolveRows));
    }
    return nblk_max <= P.cap_nblk;
}

static int plan_setup(CpPlan &P, int order, const long *dims, long R, float *const *factors, int max_hist) {
    plan_bind(P, order, dims, R, factors);
    long nblk_max = 0;
    P.cap_R = R;
    for (int m = 0; m < order; ++m) {
        P.cap_panel[m] = std::max(panel_floats(dims[m], P.Rp), dims[m] * R);
        TDX_HIP(hipMalloc(&P.panel[m], sizeof(float) * P.cap_panel[m]));
        TDX_HIP(hipMalloc(&P.G[m], sizeof(double) * R * R));
        TDX_HIP(hipMalloc(&P.Vinv[m], sizeof(double) * R * R));
        long nb = ceil_div(dims[m], kSolveRows);
        if (nb > nblk_max) nblk_max = nb;
    }
    P.cap_nblk = nblk_max;
    TDX_HIP(hipMalloc(&P.ip_part, sizeof(double) * nblk_max));
    TDX_HIP(hipMalloc(&P.gram_part, sizeof(double) * nblk_max * R * R));
    P.max_sweeps_hist = max_hist < 1 ? 1 : max_hist;
    TDX_HIP(hipMalloc(&P.fit_hist, sizeof(double) * P.max_sweeps_hist));
    TDX_HIP(hipMalloc(&P.counter, sizeof(int)));
    TDX_HIP(hipMalloc(&P.status, sizeof(int)));
    TDX_HIP(hipMalloc(&P.normx2, sizeof(double)));
    TDX_HIP(hipMalloc(&P.scratch, sizeof(double) * 4096));
    TDX_HIP(hipMalloc(&P.vinv_ws, sizeof(double) * 2 * R * R));
    TDX_HIP(hipMalloc(&P.hpart, sizeof(double) * ceil_div(R * R, 256L)));
    TDX_HIP(hipMalloc(&P.fit_done, sizeof(int)));
    TDX_HIP(hipMemset(P.fit_done, 0, sizeof(int)));
    if (P.ls_mode) {
        auto alloc = [&](void **p, size_t bytes) -> int {
            TDX_HIP(hipMalloc(p, bytes));
            P.owned.push_back(*p);
            return TDX_OK;
        };
        int rc = 0;
        for (int m = 0; m < order && !rc; ++m) {
            rc = alloc((void **)&P.U1[m], sizeof(float) * P.cap_panel[m]);
            if (!rc) rc = alloc((void **)&P.Ua[m], sizeof(float) * P.cap_panel[m]);
            if (!rc) rc = alloc((void **)&P.panel_a[m], sizeof(float) * P.cap_panel[m]);
            if (!rc) rc = alloc((void **)&P.Ga[m], sizeof(double) * R * R);
        }
        if (!rc) rc = alloc((void **)&P.ls, sizeof(LsState));
        if (!rc) rc = alloc((void **)&P.ip_part_a, sizeof(double) * nblk_max);
        if (!rc) rc = alloc((void **)&P.feat_hist, sizeof(float) * P.max_sweeps_hist * TDX_LS_FEATURES);
        if (!rc) rc = alloc((void **)&P.alpha_hist, sizeof(float) * P.max_sweeps_hist);
        if (!rc) rc = alloc((void **)&P.accepted_hist, sizeof(int) * P.max_sweeps_hist);
        if (!rc) rc = alloc((void **)&P.plain_hist, sizeof(double) * P.max_sweeps_hist);
        if (!rc && P.ngrid > 0) rc = alloc((void **)&P.grid_hist, sizeof(double) * P.max_sweeps_hist * P.ngrid);
        if (rc) return rc;
    }
    TDX_HIP(hipStreamCreateWithFlags(&P.side, hipStreamNonBlocking));
    for (int m = 0; m < 4; ++m) {
        TDX_HIP(hipEventCreateWithFlags(&P.ev_fork[m], hipEventDisableTiming));
        TDX_HIP(hipEventCreateWithFlags(&P.ev_join[m], hipEventDisableTiming));
    }
    P.events = true;
    return TDX_OK;
}

static int ensure_partial(CpPlan &P, const Storage &S) {
    size_t need = 0;
    for (int m = 0; m < P.order; ++m)
        if (S.mp[m].partial_floats() > need) need = S.mp[m].partial_floats();
    if (S.dt.on) {
        if (S.dt.partial_floats() > need) need = S.dt.partial_floats();
        if (S.dt.t_bytes() > P.t_bytes) {
            if (P.T) TDX_HIP(hipFree(P.T));
            P.T = nullptr;
            TDX_HIP(hipMalloc(&P.T, S.dt.t_bytes()));
            P.t_bytes = S.dt.t_bytes();
        }
    }
    if (need > P.partial_floats) {
        if (P.partial) TDX_HIP(hipFree(P.partial));
        P.partial = nullptr;
        TDX_HIP(hipMalloc(&P.partial, sizeof(float) * need));
        P.partial_floats = need;
    }
    size_t mneed = 0;
    for (int m = 0; m < P.order; ++m)
        if ((size_t)S.mp[m].rows_pad * P.Rp > mneed) mneed = (size_t)S.mp[m].rows_pad * P.Rp;
    if (mneed > P.msum_doubles) {
        if (P.msum) TDX_HIP(hipFree(P.msum));
        P.msum = nullptr;
        TDX_HIP(hipMalloc(&P.msum, sizeof(double) * mneed));
        P.msum_doubles = mneed;
    }
    return TDX_OK;
}

// split: BF16 storage streamed with the hi + lo Khatri-Rao block (no shared pass); *ok = false
// (and nothing set up) when some mode's plan is not the LDS kernel that implements it
static int storage_setup(CpPlan &P, Storage &S, const void *X, int dtype, const float *scales, bool split = false,
                         bool *ok = nullptr) {
    S.X = X;
    S.dtype = dtype;
    S.scales = scales;
    for (int m = 0; m < P.order; ++m) {
        int st = mttkrp_plan(S.mp[m], P.order, P.dims, m, dtype, P.R, X, 0, split);
        if (st) return st;
        if (split) {
            if (dtype != TDX_BF16 || !S.mp[m].lds) {
                if (ok) *ok = false;
                return TDX_OK;
            }
            S.mp[m].split = true;
        }
    }
    if (ok) *ok = true;
    int st = dtree_plan(S.dt, P.order, P.dims, dtype, P.R, P.shared_pass && !split);
    if (st) return st;
    return ensure_partial(P, S);
}

// (re)packs panels and Grams from the current factors
static int refresh_factors(CpPlan &P, const Storage &S, hipStream_t s) {
    // partials of a previous storage phase are never reused across a storage change
    if (P.ls_mode) k_set_int<<<1, 1, 0, s>>>(&P.ls->skip, 0);
    for (int m = 0; m < P.order; ++m) {
        const float *sc = (S.dtype == TDX_FP8_E4M3 && m == 0) ? S.scales : nullptr;
        int st = pack_panel(P.U[m], P.dims[m], P.R, P.Rp, sc, P.panel[m], s);
        if (st) return st;
        st = gram_of(P.U[m], P.dims[m], (int)P.R, P.gram_part, P.G[m], s);
        if (st) return st;
    }
    return TDX_OK;
}

// ||X||^2 of the FP32 tensor into *out; with Xbf (non-null) the same pass writes its BF16 copy.
// scratch holds 4096 doubles.  With amax (non-null, zeroed, I entries) the pass also takes the
// mode-1 slice maxima when a grid of whole slice periods exists (returns 1 in *amax_done);
// otherwise the caller computes them.
static int norm2(const void *X, int dtype, const float *scales, long I, long n, double *scratch, double *out,
                 hipStream_t s, uint16_t *Xbf = nullptr, uint32_t *amax = nullptr, int *amax_done = nullptr) {
    if (dtype != TDX_F32) return record(TDX_ERR_UNSUPPORTED);
    (void)scales;
    if (amax_done) *amax_done = 0;
    const bool aligned = ((((uintptr_t)X) | ((uintptr_t)Xbf)) & 15) == 0;
    if (aligned) {
        int blocks = std::min(4096, std::max(1, device_cus() * 8));
        bool fuse = false;
        if (amax && I % 4 == 0 && n % I == 0) {
            // grid stride 1024 blocks elements: a multiple of I for blocks a multiple of
            // I / gcd(I, 1024)
            long g = I, h = 1024;
            while (h) {
                const long t = g % h;
                g = h;
                h = t;
            }
            const long per = I / g;
            if (per <= 4096) {
                blocks = (int)(std::max(1L, blocks / per) * per);
                fuse = blocks <= 4096;
            }
        }
        if (fuse) {
            if (Xbf) k_norm2_stream<true, true><<<blocks, 256, 0, s>>>((const float *)X, n, Xbf, scratch, amax, I);
            else k_norm2_stream<false, true><<<blocks, 256, 0, s>>>((const float *)X, n, nullptr, scratch, amax, I);
            if (amax_done) *amax_done = 1;
        } else if (Xbf) {
            k_norm2_stream<true><<<blocks, 256, 0, s>>>((const float *)X, n, Xbf, scratch);
        } else {
            k_norm2_stream<false><<<blocks, 256, 0, s>>>((const float *)X, n, nullptr, scratch);
        }
        k_sum_partials<<<1, 256, 0, s>>>(scratch, blocks, out);
        TDX_LAUNCH_CHECK();
        return TDX_OK;
    }
    const int blocks = 1024;
    k_norm2_partial<float><<<blocks, 256, 0, s>>>((const float *)X, n, nullptr, I, scratch);
    k_sum_partials<<<1, 256, 0, s>>>(scratch, blocks, out);
    TDX_LAUNCH_CHECK();
    if (Xbf) return tdx_quantize_bf16((const float *)X, Xbf, n);
    return TDX_OK;
}

// enqueue one sweep (all modes) on stream s
static int enqueue_ls(CpPlan &P, const Storage &S, hipStream_t s);

static int enqueue_sweep(CpPlan &P, const Storage &S, hipStream_t s, bool fit) {
    const int R = (int)P.R;
    if (P.ls_mode) {
        k_copy_set<<<256, 256, 0, s>>>(P.set_of(P.U), P.set_of(P.U1));
        TDX_LAUNCH_CHECK();
    }
    // V^{-1} (one workgroup) and the MTTKRP depend only on the previous mode's update.
    // When the MTTKRP grid leaves a workgroup slot free on every CU (one workgroup per CU)
    // the inverse runs beside it on the side stream; otherwise ahead of it on the same
    // stream, since it would hold a CU the MTTKRP grid needs.
    // Shared pass (order 3, BF16/FP8): mode 1 from one pass over X that also writes T,
    // mode 2 from T.  The inverse (4 KB of LDS) fits beside either kernel's workgroups.
    // The fork and join between the streams cost ~15 us of replay time per mode, more
    // than the inverse itself at R <= 16 (~7 us), so small ranks keep it in line.
    // Order 3 with the shared pass, R > 16: mode 2's inverse runs inside the pass over T
    // (k_m2_vinv), after mode 1's Gram sum in line, so mode 2 has no inverse launch of its own.
    const bool fuse2 = S.dt.on && P.order == 3 && R > 16;
    auto beside_of = [&](int n) {
        if (fuse2 && n == 2) return false;
        const bool dt = S.dt.on && n > 0;
        return R > 16 && (dt || S.mp[n].rowtiles * (long)S.mp[n].nsplit <= device_cus());
    };
    // gmode >= 0: the inverse first sums mode gmode's Gram partials (gnblk blocks) into G_gmode
    auto launch_vinv = [&](int n, hipStream_t vs, int gmode = -1, int gnblk = 0) {
        const double *g2 = P.order > 2 ? P.G[2] : P.G[0], *g3 = P.order > 3 ? P.G[3] : P.G[0];
        const double *gp = gmode >= 0 ? P.gram_part : nullptr;
        double *gs = gmode >= 0 ? P.G[gmode] : nullptr;
        if (R <= 16) k_vinv<4><<<1, 256, 0, vs>>>(P.G[0], P.G[1], g2, g3, P.order, n, R, P.Vinv[n], P.status,
                                                P.vinv_ws, gp, gnblk, gs);
        else if (R <= 32) k_vinv<8><<<1, 256, 0, vs>>>(P.G[0], P.G[1], g2, g3, P.order, n, R, P.Vinv[n], P.status,
                                                     P.vinv_ws, gp, gnblk, gs);
        else k_vinv<16><<<1, 256, 0, vs>>>(P.G[0], P.G[1], g2, g3, P.order, n, R, P.Vinv[n], P.status,
                                         P.vinv_ws, gp, gnblk, gs);
    };
    // R <= 16 (every inverse in line): mode n's Gram is summed by mode n + 1's inverse
    int pending_gram = -1, pending_nblk = 0;
    int prev_nblk = 0;  // Gram-partial blocks of the mode whose Gram a forked branch sums
    // forked[n]: the fork point of mode n's inverse was recorded after the previous mode's
    // solve (its branch also sums that mode's Gram).  A side branch is enqueued after the main
    // kernel it runs beside, with the fork event recorded before that kernel: the dependencies
    // are the same, and a captured graph then has the main kernel as the first successor of
    // the fork point, which keeps the chain MTTKRP -> solve -> MTTKRP on one hardware queue
    // (with the side branch first, each of those edges crossed queues: ~9-15 us each instead
    // of ~5 at 1024^3, R 64).
    bool forked[4] = {false, false, false, false};
    for (int n = 0; n < P.order; ++n) {
        const bool dt = S.dt.on && n > 0;
        const bool beside = beside_of(n);
        if (!forked[n] && !(fuse2 && n == 2)) {
            if (beside) {
                TDX_HIP(hipEventRecord(P.ev_fork[n], s));
            } else {
                launch_vinv(n, s, pending_gram, pending_nblk);
                TDX_LAUNCH_CHECK();
                pending_gram = -1;
            }
        }
        const float *panels[4] = {P.panel[0], P.panel[1], P.panel[2], P.panel[3]};
        const float *row_scale = (S.dtype == TDX_FP8_E4M3 && n == 0) ? S.scales : nullptr;
        const int *skip = (P.ls_mode && n == 0) ? &P.ls->skip : nullptr;
        int st;
        int nsplit = S.mp[n].nsplit;
        long rows_pad = S.mp[n].rows_pad;
        if (dt && n == 1) {
            st = dtree_m1_launch(S.dt, S.X, P.panel[0], P.U[2], P.T, P.partial, s);
            nsplit = S.dt.nsplit1;
            rows_pad = S.dt.I1;
        } else if (dt && n == 2) {
            if (fuse2)
                st = dtree_m2_vinv_launch(S.dt, P.T, P.panel[1], P.partial, P.G[0], P.G[1], R, P.Vinv[2],
                                          P.status, P.vinv_ws, s);
            else
                st = dtree_m2_launch(S.dt, P.T, P.panel[1], P.partial, s);
            nsplit = S.dt.nsplit2;
            rows_pad = S.dt.I2;
        } else {
            st = mttkrp_launch(S.mp[n], S.X, panels, row_scale, P.partial, s, skip);
        }
        if (st) return st;
        if (beside) {
            // the side branch behind the fork point: (the previous mode's Gram sum,) the inverse
            TDX_HIP(hipStreamWaitEvent(P.side, P.ev_fork[n], 0));
            if (forked[n]) {
                k_gram_psum<<<(int)ceil_div((long)R * R, 256L), 256, 0, P.side>>>(P.gram_part, prev_nblk, R * R,
                                                                                 P.G[n - 1]);
                TDX_LAUNCH_CHECK();
            }
            launch_vinv(n, P.side);
            TDX_LAUNCH_CHECK();
            TDX_HIP(hipEventRecord(P.ev_join[n], P.side));
            TDX_HIP(hipStreamWaitEvent(s, P.ev_join[n], 0));
        }
        const long In = P.dims[n];
        const int nblk = (int)ceil_div(In, kSolveRows);
        const float *pscale = (S.dtype == TDX_FP8_E4M3 && n == 0) ? S.scales : nullptr;
        // with the shared pass, A is kept BF16-exact (the pass multiplies by it in BF16)
        const int quant = (S.dt.on && n == 0) ? 1 : 0;
        k_solve<<<nblk, 256, 0, s>>>(P.partial, nsplit, rows_pad, P.Rp, In, R, P.Vinv[n], P.U[n], P.panel[n], pscale,
                                     P.ip_part, P.gram_part, skip ? &P.ls->skip : nullptr, quant);
        // G_n; after the last mode together with the fit of the sweep.  The next mode's MTTKRP
        // needs only this mode's factor, so when its inverse runs beside it, G_n is summed on
        // the side stream ahead of that inverse, off the MTTKRP's path (the next solve, which
        // rewrites gram_part, waits for that inverse).  Not before the pass over T: it is as
        // short as the inverse, and launched together with it the inverse ran 76 us instead
        // of ~51 (1024^3, R 64) and became the longer branch.
        const bool early = n + 1 < P.order && beside_of(n + 1) && !(S.dt.on && n + 1 == 2);
        {
            const int gb = (int)ceil_div((long)R * R, 256L);
            if (fit && n == P.order - 1) {
                k_gram_psum_fit<<<gb, 256, 0, s>>>(P.gram_part, nblk, R, P.order, n, P.G[n], P.G[0], P.G[1],
                                                   P.order > 2 ? P.G[2] : P.G[0], P.order > 3 ? P.G[3] : P.G[0],
                                                   P.ip_part, nblk, P.normx2, P.fit_hist, P.counter, P.hpart,
                                                   P.fit_done);
            } else if (early) {
                TDX_HIP(hipEventRecord(P.ev_fork[n + 1], s));  // the branch follows mode n + 1's kernel
                forked[n + 1] = true;
                prev_nblk = nblk;
            } else if (R <= 16 && n + 1 < P.order && !beside_of(n + 1)) {
                pending_gram = n;  // the next inverse sums it (one launch fewer per mode)
                pending_nblk = nblk;
            } else {
                k_gram_psum<<<gb, 256, 0, s>>>(P.gram_part, nblk, R * R, P.G[n]);
            }
        }
        TDX_LAUNCH_CHECK();
    }
    if (P.ls_mode) return enqueue_ls(P, S, s);
    return TDX_OK;
}

// one candidate U_a (or one per grid point): features/alpha, U_a and its panels and
// Grams, mode-0 MTTKRP of U_a, exact relative error, decision, commit
static int enqueue_ls(CpPlan &P, const Storage &S, hipStream_t s) {
    const int R = (int)P.R;
    const int ncand = (P.ls_mode == 3) ? P.ngrid : 1;
    const float *sc0 = (S.dtype == TDX_FP8_E4M3) ? S.scales : nullptr;
    const long I0 = P.dims[0];
    const int nblk0 = (int)ceil_div(I0, kSolveRows);
    for (int g = 0; g < ncand; ++g) {
        const float ga = (P.ls_mode == 3) ? (float)P.grid[g] : 1.0f;
        k_ls_features<<<1, 256, 0, s>>>(P.set_of(P.U), P.set_of(P.U1), P.fit_hist, P.counter, P.ls, P.ls_mode, ga,
                                        P.mlp, P.mlp_width, P.feat_hist, P.alpha_hist, P.plain_hist);
        k_extrapolate<<<256, 256, 0, s>>>(P.set_of(P.U), P.set_of(P.U1), P.set_of(P.Ua), P.panel_a[0], P.panel_a[1],
                                          P.panel_a[2], P.panel_a[3], P.Rp, P.ls, sc0);
        for (int m = 0; m < P.order; ++m) {
            int st = gram_of(P.Ua[m], P.dims[m], R, P.gram_part, P.Ga[m], s);
            if (st) return st;
        }
        const float *pa[4] = {P.panel_a[0], P.panel_a[1], P.panel_a[2], P.panel_a[3]};
        const float *row_scale = (S.dtype == TDX_FP8_E4M3) ? S.scales : nullptr;
        int st = mttkrp_launch(S.mp[0], S.X, pa, row_scale, P.partial, s, nullptr);
        if (st) return st;
        const long mel0 = (long)S.mp[0].rows_pad * P.Rp;
        k_split_sum<<<(int)ceil_div(mel0, 256L), 256, 0, s>>>(P.partial, S.mp[0].nsplit, mel0, P.msum);
        k_ls_ip<<<nblk0, 256, 0, s>>>(P.msum, P.Rp, I0, R, P.Ua[0], P.ip_part_a);
        const double *G3 = P.order > 3 ? P.Ga[3] : P.Ga[0];
        k_ls_decide<<<1, 256, 0, s>>>(P.ip_part_a, nblk0, P.Ga[0], P.Ga[1], P.Ga[2], G3, P.order, R, P.normx2,
                                      P.fit_hist, P.counter, P.ls_mode, g, P.ngrid, P.grid_hist, P.accepted_hist,
                                      P.ls);
        if (P.ls_mode != 3)
            k_ls_commit<<<64, 256, 0, s>>>(P.set_of(P.U), P.set_of(P.Ua), P.panel[0], P.panel[1], P.panel[2],
                                           P.panel[3], P.panel_a[0], P.panel_a[1], P.panel_a[2], P.panel_a[3], P.Rp,
                                           P.G[0], P.G[1], P.G[2], P.G[3], P.Ga[0], P.Ga[1], P.Ga[2], P.Ga[3], P.ls);
        TDX_LAUNCH_CHECK();
    }
    return TDX_OK;
}

struct SweepGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    ~SweepGraph() {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
    }
};

static int capture_sweep(CpPlan &P, const Storage &S, hipStream_t s, bool fit, SweepGraph &G) {
    TDX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int st = enqueue_sweep(P, S, s, fit);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return record_hip(e, "hipStreamEndCapture", __FILE__, __LINE__);
    G.g = g;
    TDX_HIP(hipGraphInstantiate(&G.ge, g, nullptr, nullptr, 0));
    return TDX_OK;
}

// Runs `sweeps` sweeps on storage S; records per-sweep ms if sweep_ms != NULL, and the device
// time from the event `origin` to the end of every sweep if elapsed_ms != NULL.
static int run_sweeps(CpPlan &P, const Storage &S, hipStream_t s, int sweeps, bool fit, bool graph, float *sweep_ms,
                      double tol, double target, int *done, hipEvent_t origin = nullptr,
                      float *elapsed_ms = nullptr) {
    *done = 0;
    if (sweeps <= 0) return TDX_OK;
    SweepGraph G;
    if (graph) {
        int st = capture_sweep(P, S, s, fit, G);
        if (st) return st;
    }
    if (!origin) elapsed_ms = nullptr;
    std::vector<hipEvent_t> evs;
    if (sweep_ms || elapsed_ms) {
        // timing only (read after hipEventSynchronize of the last one; results reach the host
        // through the caller's stream synchronization): no system-scope cache writeback and
        // invalidation at each record, which cost ~6 us between sweeps at 1024^3
        evs.resize(sweeps + 1);
        for (auto &e : evs) TDX_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        TDX_HIP(hipEventRecord(evs[0], s));
    }
    double prev = -1.0;
    int st = TDX_OK;
    for (int it = 0; it < sweeps; ++it) {
        if (graph) {
            hipError_t e = hipGraphLaunch(G.ge, s);
            if (e != hipSuccess) { st = record_hip(e, "hipGraphLaunch", __FILE__, __LINE__); break; }
        } else {
            st = enqueue_sweep(P, S, s, fit);
            if (st) break;
        }
        if (!evs.empty()) {
            hipError_t e = hipEventRecord(evs[it + 1], s);
            if (e != hipSuccess) { st = record_hip(e, "hipEventRecord", __FILE__, __LINE__); break; }
        }
        *done = it + 1;
        if ((tol > 0.0 || target > 0.0) && fit) {
            int c = 0;
            double f = 0.0;
            hipError_t e = hipStreamSynchronize(s);
            if (e == hipSuccess) e = hipMemcpy(&c, P.counter, sizeof(int), hipMemcpyDeviceToHost);
            if (e == hipSuccess && c > 0) e = hipMemcpy(&f, P.fit_hist + (c - 1), sizeof(double), hipMemcpyDeviceToHost);
            if (e != hipSuccess) { st = record_hip(e, "fit readback", __FILE__, __LINE__); break; }
            if (tol > 0.0 && prev >= 0.0 && fabs(prev - f) < tol) break;
            if (target > 0.0 && f <= target) break;
            prev = f;
        }
    }
    if (!evs.empty() && st == TDX_OK) {
        TDX_HIP(hipEventSynchronize(evs[*done]));
        for (int it = 0; it < *done; ++it) {
            if (sweep_ms) TDX_HIP(hipEventElapsedTime(&sweep_ms[it], evs[it], evs[it + 1]));
            if (elapsed_ms) TDX_HIP(hipEventElapsedTime(&elapsed_ms[it], origin, evs[it + 1]));
        }
    }
    for (auto &e : evs) (void)hipEventDestroy(e);
    return st;
}

// ---- plan cache: the device buffers of a CP-ALS call are kept for the next call of the same
// shape (a loop over many tensors of one shape then allocates nothing after the first call).
// A plan is taken out of the cache while a call uses it, so concurrent calls never share one.
struct PlanKey {
    int ls_mode = 0, ngrid = 0, mlp_width = 0;
    bool operator==(const PlanKey &o) const {
        return ls_mode == o.ls_mode && ngrid == o.ngrid && mlp_width == o.mlp_width;
    }
};
static constexpr size_t kPlanCacheSize = 4;
static std::mutex g_plan_mutex;
static std::vector<std::pair<PlanKey, std::unique_ptr<CpPlan>>> g_plans;  // most recent first

// a cached plan of the same line-search configuration whose buffers hold the shape
static std::unique_ptr<CpPlan> take_plan(const PlanKey &k, int order, const long *dims, long R, int max_hist) {
    std::lock_guard<std::mutex> lk(g_plan_mutex);
    for (size_t i = 0; i < g_plans.size(); ++i)
        if (g_plans[i].first == k && plan_fits(*g_plans[i].second, order, dims, R, max_hist)) {
            std::unique_ptr<CpPlan> p = std::move(g_plans[i].second);
            g_plans.erase(g_plans.begin() + (long)i);
            return p;
        }
    return nullptr;
}

static void give_plan(const PlanKey &k, std::unique_ptr<CpPlan> p) {
    std::unique_ptr<CpPlan> evicted;
    {
        std::lock_guard<std::mutex> lk(g_plan_mutex);
        g_plans.insert(g_plans.begin(), std::make_pair(k, std::move(p)));
        if (g_plans.size() > kPlanCacheSize) {
            evicted = std::move(g_plans.back().second);
            g_plans.pop_back();
        }
    }
    // evicted (if any) is released here, outside the lock
}

void release_cp_plans() {
    std::vector<std::pair<PlanKey, std::unique_ptr<CpPlan>>> old;
    {
        std::lock_guard<std::mutex> lk(g_plan_mutex);
        old.swap(g_plans);
    }
}

}  // namespace tdx

using namespace tdx;

extern "C" {

void tdx_cp_default_options(tdx_cp_options *o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->max_sweeps = 500;
    o->refine_sweeps = 0;
    o->storage = TDX_F32;
    o->init = 0;
    o->seed = 1234;
    o->tol = 0.0;
    o->record_fit = 1;
    o->line_search = 0;
    o->use_graph = 1;
    o->mlp_width = 64;
    o->shared_pass = 1;
}

int tdx_cp_als_ex(const float *X, int order, const long *dims, long rank, float *const *factors,
                  const tdx_cp_options *opt_in, tdx_cp_report *report, double *fit_hist, float *sweep_ms) {
    if (!X || !dims || !factors || order < 3 || order > 4 || rank <= 0) return record(TDX_ERR_ARG);
    if (rank > kMaxR) return record(TDX_ERR_UNSUPPORTED);
    tdx_cp_options opt;
    if (opt_in) opt = *opt_in;
    else tdx_cp_default_options(&opt);
    if (opt.max_sweeps < 0 || opt.refine_sweeps < 0 || opt.opening_sweeps < 0) return record(TDX_ERR_ARG);
    if (opt.storage != TDX_F32 && opt.storage != TDX_BF16 && opt.storage != TDX_FP8_E4M3) return record(TDX_ERR_ARG);
    // the extrapolation step runs on FP32 storage only: with BF16 / FP8 storage it is switched
    // off (a run of it ended in a host segmentation fault whose cause is not established)
    if (opt.line_search != 0 && opt.storage != TDX_F32) return record(TDX_ERR_UNSUPPORTED);
    long N = 1;
    for (int m = 0; m < order; ++m) {
        if (dims[m] <= 0 || !factors[m]) return record(TDX_ERR_ARG);
        N *= dims[m];
    }
    // Graph capture is not allowed on the legacy default stream: when the library stream
    // is the default one, run on a private blocking stream (it orders after earlier
    // default-stream work by the legacy-stream rules) and join it before returning.
    hipStream_t s = stream();
    hipStream_t owned = nullptr;
    if (!s) {
        TDX_HIP(hipStreamCreate(&owned));
        s = owned;
    }
    struct StreamGuard {
        hipStream_t st;
        ~StreamGuard() {
            if (st) {
                (void)hipStreamSynchronize(st);
                (void)hipStreamDestroy(st);
            }
        }
    } guard{owned};
    const int total = opt.max_sweeps + (opt.storage == TDX_F32 ? 0 : opt.refine_sweeps);
    if (opt.line_search < 0 || opt.line_search > 3) return record(TDX_ERR_ARG);
    if (opt.line_search == 1 && (!opt.mlp_params || opt.mlp_width <= 0 || opt.mlp_width > 128))
        return record(TDX_ERR_ARG);
    if (opt.line_search == 3 && (opt.ngrid <= 0 || opt.ngrid > 8)) return record(TDX_ERR_ARG);
    // the plan (device buffers, side stream, events) of an earlier call with the same shape is
    // reused; it returns to the cache only when this call succeeds
    PlanKey key{};
    key.ls_mode = opt.line_search;
    key.ngrid = opt.line_search == 3 ? opt.ngrid : 0;
    key.mlp_width = opt.line_search == 1 ? opt.mlp_width : 64;
    std::unique_ptr<CpPlan> Pp = take_plan(key, order, dims, rank, total);
    const bool fresh = !Pp;
    if (fresh) Pp.reset(new CpPlan);
    CpPlan &P = *Pp;
    P.ls_mode = opt.line_search;
    P.shared_pass = opt.shared_pass != 0;
    if (P.ls_mode == 1) {
        P.mlp = opt.mlp_params;
        P.mlp_width = opt.mlp_width;
    }
    if (P.ls_mode == 3) {
        P.ngrid = opt.ngrid;
        for (int g = 0; g < opt.ngrid; ++g) P.grid[g] = opt.grid[g];
    }
    if (P.ls_mode) opt.record_fit = 1;
    int st = TDX_OK;
    if (fresh) {
        st = plan_setup(P, order, dims, rank, factors, total);
        if (st) return st;
    } else {
        plan_bind(P, order, dims, rank, factors);
    }
    if (P.ls_mode) k_ls_init<<<1, 1, 0, s>>>(P.ls);
    // initialisation: modes 1..d-1 from the seeded generator unless the caller supplied them
    if (opt.init == 0) {
        for (int m = 1; m < order; ++m) {
            st = tdx_fill_uniform(factors[m], dims[m] * rank, opt.seed, (uint32_t)m);
            if (st) return st;
        }
    }
    // the timed region starts before the first device work: the norm of X, the conversion
    // of X to the main-phase storage and the factor panels are part of it
    struct EventGuard {
        hipEvent_t e[3] = {nullptr, nullptr, nullptr};
        ~EventGuard() {
            for (auto &x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evg;
    for (auto &x : evg.e) TDX_HIP(hipEventCreate(&x));
    (void)hipEventRecord(evg.e[0], s);
    k_set_int<<<1, 1, 0, s>>>(P.counter, 0);
    k_set_int<<<1, 1, 0, s>>>(P.status, 0);
    // storage copies for the main phase (workspace pool: no allocation after the first call);
    // the BF16 copy is written by the pass that sums ||X||^2
    // BF16 / FP8 storage: the first `opening` of the max_sweeps main sweeps read the FP32 tensor,
    // or (opening_split) a BF16 copy with the split Khatri-Rao block
    const int opening = (opt.storage == TDX_F32) ? 0 : std::min(opt.opening_sweeps, opt.max_sweeps);
    const bool split_open = opening > 0 && opt.opening_split && opt.storage != TDX_F32;
    void *Xq = nullptr, *Xb = nullptr;  // main-phase copy; BF16 copy (= Xq for BF16 storage)
    float *scales = nullptr;
    uint32_t *amax = nullptr;
    if (opt.storage == TDX_BF16) {
        st = workspace(WS_STORE, sizeof(uint16_t) * N, &Xq);
        if (st) return st;
        Xb = Xq;
    } else if (opt.storage == TDX_FP8_E4M3) {
        void *sp = nullptr, *ap = nullptr;
        st = workspace(WS_STORE, N, &Xq);
        if (!st) st = workspace(WS_STORE_SCALES, sizeof(float) * dims[0], &sp);
        if (!st) st = workspace(WS_MISC3, sizeof(uint32_t) * dims[0], &ap);
        if (!st && split_open) st = workspace(WS_STORE_BF16, sizeof(uint16_t) * N, &Xb);
        if (st) return st;
        scales = (float *)sp;
        amax = (uint32_t *)ap;
        k_fill_u32_dev<<<1, 256, 0, s>>>(amax, dims[0], 0u);
    }
    // one pass: ||X||^2, the BF16 copy (if any) and, for FP8 storage, the slice maxima when the
    // tensor shape allows (otherwise a pass of their own)
    int amax_done = 0;
    st = norm2(X, TDX_F32, nullptr, dims[0], N, P.scratch, P.normx2, s, (uint16_t *)Xb, amax, &amax_done);
    if (st) return st;
    Storage S32;
    st = storage_setup(P, S32, X, TDX_F32, nullptr);
    if (st) return st;
    Storage Smain = S32;
    if (opt.storage == TDX_BF16) {
        st = storage_setup(P, Smain, Xq, TDX_BF16, nullptr);
    } else if (opt.storage == TDX_FP8_E4M3) {
        if (!amax_done) st = slice_amax(X, amax, dims[0], N / dims[0], s);
        if (!st) st = fp8_from_amax(X, amax, (uint8_t *)Xq, scales, dims[0], N / dims[0], s);
        if (!st) st = storage_setup(P, Smain, Xq, TDX_FP8_E4M3, scales);
    }
    if (st) return record(st);
    Storage Sopen = S32;
    if (split_open) {
        Storage Ss;
        bool ok = false;
        st = storage_setup(P, Ss, Xb, TDX_BF16, nullptr, true, &ok);
        if (st) return record(st);
        if (ok) Sopen = Ss;
    }
    float *elapsed = opt.trace ? opt.trace->elapsed_ms : nullptr;
    int done_open = 0, done_main = 0, done_ref = 0;
    st = refresh_factors(P, opening ? Sopen : Smain, s);
    (void)hipEventRecord(evg.e[1], s);
    if (!st && opening > 0) {
        st = run_sweeps(P, Sopen, s, opening, opt.record_fit != 0, opt.use_graph != 0, sweep_ms, 0.0, 0.0, &done_open,
                        evg.e[0], elapsed);
        if (!st) st = refresh_factors(P, Smain, s);
    }
    if (!st)
        st = run_sweeps(P, Smain, s, opt.max_sweeps - opening, opt.record_fit != 0, opt.use_graph != 0,
                        sweep_ms ? sweep_ms + done_open : nullptr, opt.tol,
                        opt.storage == TDX_F32 ? opt.target_relerr : 0.0, &done_main, evg.e[0],
                        elapsed ? elapsed + done_open : nullptr);
    done_main += done_open;
    if (!st && opt.storage != TDX_F32 && opt.refine_sweeps > 0) {
        // FP32 refinement: the mode-0 panel must be unscaled again (rebuilt by the first solve)
        st = refresh_factors(P, S32, s);
        if (!st)
            st = run_sweeps(P, S32, s, opt.refine_sweeps, opt.record_fit != 0, opt.use_graph != 0,
                            sweep_ms ? sweep_ms + done_main : nullptr, 0.0, 0.0, &done_ref, evg.e[0],
                            elapsed ? elapsed + done_main : nullptr);
    }
    (void)hipEventRecord(evg.e[2], s);
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess && !st) st = record_hip(e, "hipStreamSynchronize", __FILE__, __LINE__);
    float ms = 0.0f, setup_ms = 0.0f;
    (void)hipEventElapsedTime(&ms, evg.e[0], evg.e[2]);
    (void)hipEventElapsedTime(&setup_ms, evg.e[0], evg.e[1]);
    if (st) return record(st);
    int status = 0, count = 0;
    TDX_HIP(hipMemcpy(&status, P.status, sizeof(int), hipMemcpyDeviceToHost));
    TDX_HIP(hipMemcpy(&count, P.counter, sizeof(int), hipMemcpyDeviceToHost));
    double last = 0.0;
    if (count > 0) {
        std::vector<double> h(count);
        TDX_HIP(hipMemcpy(h.data(), P.fit_hist, sizeof(double) * count, hipMemcpyDeviceToHost));
        if (fit_hist) memcpy(fit_hist, h.data(), sizeof(double) * count);
        last = h[count - 1];
    }
    if (opt.trace && P.ls_mode && count > 0) {
        tdx_cp_trace *tr = opt.trace;
        if (tr->alpha) TDX_HIP(hipMemcpy(tr->alpha, P.alpha_hist, sizeof(float) * count, hipMemcpyDeviceToHost));
        if (tr->accepted)
            TDX_HIP(hipMemcpy(tr->accepted, P.accepted_hist, sizeof(int) * count, hipMemcpyDeviceToHost));
        if (tr->features)
            TDX_HIP(hipMemcpy(tr->features, P.feat_hist, sizeof(float) * count * TDX_LS_FEATURES,
                              hipMemcpyDeviceToHost));
        if (tr->plain_relerr)
            TDX_HIP(hipMemcpy(tr->plain_relerr, P.plain_hist, sizeof(double) * count, hipMemcpyDeviceToHost));
        if (tr->grid_relerr && P.grid_hist)
            TDX_HIP(hipMemcpy(tr->grid_relerr, P.grid_hist, sizeof(double) * count * P.ngrid, hipMemcpyDeviceToHost));
    }
    int result = status ? TDX_WARN_SINGULAR : TDX_OK;
    if (opt.record_fit && count > 0 && !isfinite(last)) result = TDX_ERR_NONFINITE;
    give_plan(key, std::move(Pp));
    if (report) {
        report->sweeps_done = done_main + done_ref;
        report->status = result;
        report->relerr_fit = last;
        report->seconds = ms * 1e-3;
        report->setup_seconds = setup_ms * 1e-3;
    }
    return record(result);
}

int tdx_cp_als(const float *X, float *A, float *B, float *C, long a, long b, long c, long r) {
    const long dims[3] = {a, b, c};
    float *f[3] = {A, B, C};
    tdx_cp_options o;
    tdx_cp_default_options(&o);
    o.record_fit = 0;
    return tdx_cp_als_ex(X, 3, dims, r, f, &o, nullptr, nullptr, nullptr);
}

int tdx_cp_als4(const float *X, float *A, float *B, float *C, float *D, long a, long b, long c, long d, long r) {
    const long dims[4] = {a, b, c, d};
    float *f[4] = {A, B, C, D};
    tdx_cp_options o;
    tdx_cp_default_options(&o);
    o.record_fit = 0;
    return tdx_cp_als_ex(X, 4, dims, r, f, &o, nullptr, nullptr, nullptr);
}

int tdx_cp_tensorcore(const float *X, float *A, float *B, float *C, long a, long b, long c, long r) {
    const long dims[3] = {a, b, c};
    float *f[3] = {A, B, C};
    tdx_cp_options o;
    tdx_cp_default_options(&o);
    o.storage = TDX_BF16;
    o.refine_sweeps = 2;
    o.max_sweeps = 498;
    o.record_fit = 0;
    return tdx_cp_als_ex(X, 3, dims, r, f, &o, nullptr, nullptr, nullptr);
}

}  // extern "C"
