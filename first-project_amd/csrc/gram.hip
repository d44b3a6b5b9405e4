// gram.hip -- Gram of a mode-n unfolding on MFMA, G = X(n) X(n)^T, In <= 128.
//
// The unfolding is never formed (the reference builds it with tensor2matrix and, for order
// 4, identity-matrix GEMMs, src/tucker4_hosvd.cu:23-37, then calls cublasSgemm).  X is read
// as the (P, In, Q) view, column k of X(n) being the fibre (p, q).  One K step is KS
// consecutive columns; every lane loads one 16-byte vector per 16-row block, and the same
// registers serve as the A operand (rows i) and the B operand (columns j) of the MFMA, since
// B = A^T.  Only the 36 block pairs bi <= bj of the 8 x 8 block grid are multiplied.
//
//  case B (P > 1, P % KS == 0): K is contiguous along p; a lane's 16 bytes are KS/4 (FP32:
//    4, BF16: 8) consecutive p of one row, rows of block b are 16 b + 0..15.
//  case A (P == 1): rows are contiguous and K (= q) is strided; a lane loads its 8 (BF16) or
//    4 (FP32) consecutive rows at 8 (4) columns and transposes in registers, so row block t
//    holds rows 8 m + t (BF16) or 64 h + 4 m + t (FP32, block 4 h + t), m = lane % 16.
//
// The 4 waves of a workgroup take contiguous quarters of its K range and are summed in LDS
// in a fixed order; each workgroup writes one 128 x 128 FP32 partial, and the partials are
// summed in FP64 by the caller (deterministic for a given launch shape).
#include "common.hpp"

namespace tdx {

typedef float gf32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 gbf16x8 __attribute__((ext_vector_type(8)));

static constexpr int kGramTile = 128;
static constexpr int kGramLd = kGramTile + 1;

template <int ST> struct GramElem;
template <> struct GramElem<0> { typedef float T; typedef float4 V; };
template <> struct GramElem<1> { typedef uint16_t T; typedef uint4 V; };

__device__ __forceinline__ float4 gmask(float4 v, bool ok) {
    const uint32_t m = ok ? 0xffffffffu : 0u;
    return float4{__uint_as_float(__float_as_uint(v.x) & m), __uint_as_float(__float_as_uint(v.y) & m),
                  __uint_as_float(__float_as_uint(v.z) & m), __uint_as_float(__float_as_uint(v.w) & m)};
}
__device__ __forceinline__ uint4 gmask(uint4 v, bool ok) {
    const uint32_t m = ok ? 0xffffffffu : 0u;
    return uint4{v.x & m, v.y & m, v.z & m, v.w & m};
}
__device__ __forceinline__ float f4_get(const float4 &v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}
__device__ __forceinline__ uint32_t u4_at(const uint4 &v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}
__device__ __forceinline__ uint32_t h16(const uint4 &v, int t) {
    const uint32_t w = u4_at(v, t >> 1);
    return (t & 1) ? (w >> 16) : (w & 0xffffu);
}

// row of the tile held by MFMA row m of row block b
template <int ST, bool CA> __device__ __forceinline__ int gram_row(int b, int m) {
    if constexpr (!CA) return 16 * b + m;
    else if constexpr (ST == 1) return 8 * m + b;
    else return 64 * (b >> 2) + 4 * m + (b & 3);
}

template <int ST, bool CA>
__global__ __launch_bounds__(256, CA ? 1 : 2) void k_gram_mfma(const void *__restrict__ Xv, long P, long In, long Q,
                                                     long steps, long kper, float *__restrict__ part) {
    typedef typename GramElem<ST>::T ET;
    typedef typename GramElem<ST>::V VT;
    constexpr int KS = (ST == 1) ? 32 : 16;  // K columns per step
    constexpr int NPAIR = 36;
    __shared__ float red[kGramTile * kGramLd];
    const ET *__restrict__ X = reinterpret_cast<const ET *>(Xv);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lm = lane & 15, lg = lane >> 4;
    const long kbeg = (long)blockIdx.x * kper;
    const long kend = min(kbeg + kper, steps);
    const long quarter = (kend - kbeg + 3) / 4;
    const long s0 = kbeg + wave * quarter;
    const long s1 = min(s0 + quarter, kend);
    const long nst = (s0 < s1) ? s1 - s0 : 0;
    const long PB = CA ? 1 : P / KS;  // steps per q in case B

    gf32x4 acc[NPAIR];
#pragma unroll
    for (int i = 0; i < NPAIR; ++i) acc[i] = gf32x4{0.f, 0.f, 0.f, 0.f};

    struct Slot {
        VT v[8];
        uint32_t ok;
    };
    // per-lane fixed parts of the addresses
    long roff[8];  // case B: row offsets a * P; case A: the lane's first row
    bool rok[8];
    if constexpr (!CA) {
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const long a = 16 * b + lm;
            rok[b] = a < In;
            roff[b] = (rok[b] ? a : 0) * P;
        }
    } else {
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const long a0 = (ST == 1) ? 8 * lm : 64 * (h >> 2) + 4 * lm;
            rok[h] = a0 < In;
            roff[h] = rok[h] ? a0 : 0;
        }
    }
    // the wave's next step as (q, p-chunk) for case B or as the step index for case A,
    // advanced incrementally (no divisions in the loop); past the last step it stays put,
    // so the surplus issues of the pipeline tail re-read the last step (never consumed)
    long f_q = CA ? s0 : s0 / PB, f_pc = CA ? 0 : s0 % PB, issued = 0;
    auto issue = [&](Slot &sl) {
        sl.ok = 0;
        if constexpr (!CA) {
            const ET *base = X + f_pc * KS + ((ST == 1) ? 8 : 4) * lg + f_q * P * In;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                sl.v[b] = *reinterpret_cast<const VT *>(base + roff[b]);
                sl.ok |= (uint32_t)rok[b] << b;
            }
        } else if constexpr (ST == 1) {
            // 8 columns q0 + 8 lg + j, the lane's 8 rows each
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const long qq = f_q * KS + 8 * lg + j;
                const bool in = rok[0] && qq < Q;
                sl.v[j] = *reinterpret_cast<const VT *>(X + roff[0] + (in ? qq : 0) * In);
                sl.ok |= (uint32_t)in << j;
            }
        } else {
            // loads [h][c]: rows 64 h + 4 lm .. +3 at column q0 + 4 lg + c
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const long qq = f_q * KS + 4 * lg + c;
                    const bool in = rok[4 * h] && qq < Q;
                    sl.v[4 * h + c] = *reinterpret_cast<const VT *>(X + roff[4 * h] + (in ? qq : 0) * In);
                    sl.ok |= (uint32_t)in << (4 * h + c);
                }
        }
        if (++issued < nst) {
            if constexpr (CA) {
                ++f_q;
            } else if (++f_pc == PB) {
                f_pc = 0;
                ++f_q;
            }
        }
    };
    auto compute = [&](const Slot &sl) {
        if constexpr (ST == 1) {
            gbf16x8 op[8];
            if constexpr (!CA) {
#pragma unroll
                for (int b = 0; b < 8; ++b) op[b] = __builtin_bit_cast(gbf16x8, gmask(sl.v[b], (sl.ok >> b) & 1u));
            } else {
                uint4 v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = gmask(sl.v[j], (sl.ok >> j) & 1u);
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    uint4 o;
                    o.x = h16(v[0], t) | (h16(v[1], t) << 16);
                    o.y = h16(v[2], t) | (h16(v[3], t) << 16);
                    o.z = h16(v[4], t) | (h16(v[5], t) << 16);
                    o.w = h16(v[6], t) | (h16(v[7], t) << 16);
                    op[t] = __builtin_bit_cast(gbf16x8, o);
                }
            }
            int pi = 0;
#pragma unroll
            for (int bi = 0; bi < 8; ++bi)
#pragma unroll
                for (int bj = bi; bj < 8; ++bj, ++pi)
                    acc[pi] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(op[bi], op[bj], acc[pi], 0, 0, 0);
        } else {
            float4 v[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = gmask(sl.v[j], (sl.ok >> j) & 1u);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                float op[8];
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    if constexpr (!CA) op[b] = f4_get(v[b], c);               // block b, column 4 lg + c
                    else op[b] = f4_get(v[4 * (b >> 2) + c], b & 3);           // block 4h+t, column 4 lg + c
                }
                int pi = 0;
#pragma unroll
                for (int bi = 0; bi < 8; ++bi)
#pragma unroll
                    for (int bj = bi; bj < 8; ++bj, ++pi)
                        acc[pi] = __builtin_amdgcn_mfma_f32_16x16x4f32(op[bi], op[bj], acc[pi], 0, 0, 0);
            }
        }
    };

    // PD register slots, PD - 1 steps in flight per wave (case A runs one workgroup per CU
    // and keeps two steps in flight; case B runs two per CU and keeps one)
    constexpr int PD = CA ? 3 : 2;
    Slot ring[PD];
    if (nst > 0) {
#pragma unroll
        for (int p = 0; p < PD - 1; ++p) issue(ring[p]);
        for (long base = 0; base < nst; base += PD) {
#pragma unroll
            for (int p = 0; p < PD; ++p) {
                issue(ring[(p + PD - 1) % PD]);
                __builtin_amdgcn_sched_barrier(0);  // keep the prefetch ahead of the MFMAs
                if (base + p < nst) compute(ring[p]);
            }
        }
    }

    // fixed-order sum of the 4 waves into the LDS tile (both triangles)
    for (int wv = 0; wv < 4; ++wv) {
        if (wave == wv) {
            int pi = 0;
#pragma unroll
            for (int bi = 0; bi < 8; ++bi)
#pragma unroll
                for (int bj = bi; bj < 8; ++bj, ++pi) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int i = gram_row<ST, CA>(bi, 4 * lg + v), j = gram_row<ST, CA>(bj, lm);
                        const float x = acc[pi][v];
                        if (wv == 0) {
                            red[i * kGramLd + j] = x;
                            if (bi != bj) red[j * kGramLd + i] = x;
                        } else {
                            red[i * kGramLd + j] += x;
                            if (bi != bj) red[j * kGramLd + i] += x;
                        }
                    }
                }
        }
        __syncthreads();
    }
    float *dst = part + (long)blockIdx.x * kGramTile * kGramTile;
    for (int e = threadIdx.x; e < kGramTile * kGramTile; e += 256) dst[e] = red[(e >> 7) * kGramLd + (e & 127)];
}

// Partial Grams of the (P, In, Q) view: returns the number of partials written to part
// (each kGramTile^2 floats, row-major, rows/cols >= In zero), or 0 if the shape is not
// supported here (the caller falls back), or -1 on a launch error.
int gram_mfma_launch(const void *X, int xtype, long P, long In, long Q, float *part, int max_parts, hipStream_t s) {
    const int st = (xtype == TDX_BF16) ? 1 : 0;
    if (xtype != TDX_BF16 && xtype != TDX_F32) return 0;
    if (In < 1 || In > kGramTile || ((uintptr_t)X & 15)) return 0;
    const long KS = st ? 32 : 16;
    const bool ca = (P == 1);
    long steps;
    if (ca) {
        if (In % (st ? 8 : 4)) return 0;
        steps = ceil_div(Q, KS);
    } else {
        if (P % KS) return 0;
        steps = (P / KS) * Q;
    }
    if (steps < 1) return 0;
    long nblk = ceil_div(steps, 32L);  // at least 8 steps per wave
    if (nblk > max_parts) nblk = max_parts;
    const long kper = ceil_div(steps, nblk);
    nblk = ceil_div(steps, kper);
    const dim3 g((unsigned)nblk), b(256);
    if (st) {
        if (ca) k_gram_mfma<1, true><<<g, b, 0, s>>>(X, P, In, Q, steps, kper, part);
        else k_gram_mfma<1, false><<<g, b, 0, s>>>(X, P, In, Q, steps, kper, part);
    } else {
        if (ca) k_gram_mfma<0, true><<<g, b, 0, s>>>(X, P, In, Q, steps, kper, part);
        else k_gram_mfma<0, false><<<g, b, 0, s>>>(X, P, In, Q, steps, kper, part);
    }
    if (hipGetLastError() != hipSuccess) return -1;
    return (int)nblk;
}

}  // namespace tdx
