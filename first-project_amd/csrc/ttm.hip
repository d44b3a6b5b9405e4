// ttm.hip -- mode-n tensor-times-matrix on MFMA, Y = X x_n U^T:
//   Y[p, a, q] = sum_i X[p, i, q] U(i, a),  U(i, a) = U[i * ldi + a * lda],
// X the (P, In, Q) view of an FP32 or BF16 tensor, Y the (P, Ra, Q) FP32 result.  Replaces
// the cublasSgemm(Strided)Batched calls of src/ttm.cu:21,29,37 and the TTM chains of
// src/tucker_hosvd.cu:72-80, with no unfolded copy of X.
//
// A wave owns a tile of 128 output rows and all Ra <= 64 columns, with the whole K = In <= 128
// range; U is held in registers as the MFMA B operand (BF16 for BF16 storage, FP32 otherwise).
//  TR (P > 1, P % 8 == 0 for BF16, P % 4 == 0 for FP32): output rows are the fibres
//    k = p + P q, contiguous along p in X but strided along the contracted i, so a lane loads
//    16 bytes of consecutive p at each of its K columns and transposes in registers (row
//    block t holds rows 8 m + t, BF16, or 64 h + 4 m + t, FP32, m = lane % 16).
//  !TR (P == 1): output rows are q, the contracted i is contiguous: direct 16-byte loads.
// The accumulator tile goes through LDS (32 columns at a time) so that Y is written with
// contiguous 256-byte stores.  X is read with nontemporal loads (streamed once).  Each output is a sum over i in a fixed order.
#include "common.hpp"
#include <type_traits>

namespace tdx {

typedef float tf32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 tbf16x8 __attribute__((ext_vector_type(8)));

template <int ST> struct TtmElem;
template <> struct TtmElem<0> { typedef float T; typedef float4 V; };
template <> struct TtmElem<1> { typedef uint16_t T; typedef uint4 V; };

__device__ __forceinline__ float4 tmask(float4 v, bool ok) {
    const uint32_t m = ok ? 0xffffffffu : 0u;
    return float4{__uint_as_float(__float_as_uint(v.x) & m), __uint_as_float(__float_as_uint(v.y) & m),
                  __uint_as_float(__float_as_uint(v.z) & m), __uint_as_float(__float_as_uint(v.w) & m)};
}
__device__ __forceinline__ uint4 tmask(uint4 v, bool ok) {
    const uint32_t m = ok ? 0xffffffffu : 0u;
    return uint4{v.x & m, v.y & m, v.z & m, v.w & m};
}
__device__ __forceinline__ float tf4(const float4 &v, int i) { return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w; }
__device__ __forceinline__ uint32_t th16(const uint4 &v, int t) {
    const uint32_t w = (t >> 1) == 0 ? v.x : (t >> 1) == 1 ? v.y : (t >> 1) == 2 ? v.z : v.w;
    return (t & 1) ? (w >> 16) : (w & 0xffffu);
}
__device__ __forceinline__ uint32_t tpack_bf16(float lo, float hi) {
    auto rne = [](float f) -> uint32_t {
        const uint32_t u = __float_as_uint(f);
        if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
        return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    };
    return rne(lo) | (rne(hi) << 16);
}

static constexpr int kTtmMaxIn = 128;
static constexpr int kTtmCsLd = 129;  // LDS staging row pitch (floats)

template <int ST, bool TR>
__device__ __forceinline__ int ttm_row(int b, int m) {
    if constexpr (!TR) return 16 * b + m;
    else if constexpr (ST == 1) return 8 * m + b;
    else return 64 * (b >> 2) + 4 * m + (b & 3);
}

// one launch = gridDim.x workgroups of 4 waves; workgroup w takes tile groups w, w + grid, ...
// (4 tiles of 128 rows per group, one per wave, the same trip count for every wave)
template <int ST, bool TR, int NBA>
__global__ __launch_bounds__(256, 1) void k_ttm_mfma(const void *__restrict__ Xv, long P, long In, long Q,
                                                    const float *__restrict__ U, long lda, long ldi, long Ra,
                                                    float *__restrict__ Y) {
    typedef typename TtmElem<ST>::T ET;
    typedef typename TtmElem<ST>::V VT;
    constexpr int KS = (ST == 1) ? 32 : 16;        // K per step
    constexpr int NSTEP = kTtmMaxIn / KS;           // steps covering In <= 128
    __shared__ float cs[4][32 * kTtmCsLd];
    const ET *__restrict__ X = reinterpret_cast<const ET *>(Xv);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lm = lane & 15, lg = lane >> 4;
    const long rows = TR ? P * Q : Q;
    const long ntiles = ceil_div(rows, 128L);
    const long ngroups = ceil_div(ntiles, 4L);
    const int nsteps = (int)ceil_div(In, (long)KS);
    float *cw = cs[wave];

    // B operands: U(i, a) for a = 16 nb + lm at the lane's K slots of every step, read from
    // a coalesced copy of U in LDS (the staging area is free until the first tile)
    typedef typename std::conditional<ST == 1, tbf16x8, float4>::type BT;
    BT bop[NSTEP][NBA];
    {
        float *us = &cs[0][0];  // [a][i], i < 128, a < 16 NBA
        const int na = 16 * NBA;
        for (int e = threadIdx.x; e < na * kTtmMaxIn; e += 256) {
            const int i = e % kTtmMaxIn, a = e / kTtmMaxIn;
            us[e] = (i < In && a < Ra) ? U[(long)i * ldi + (long)a * lda] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NSTEP; ++s)
#pragma unroll
            for (int nb = 0; nb < NBA; ++nb) {
                const float *ua = us + (16 * nb + lm) * kTtmMaxIn + s * KS + ((ST == 1) ? 8 : 4) * lg;
                if constexpr (ST == 1) {
                    const float4 u0 = *reinterpret_cast<const float4 *>(ua);
                    const float4 u1 = *reinterpret_cast<const float4 *>(ua + 4);
                    uint4 w;
                    w.x = tpack_bf16(u0.x, u0.y);
                    w.y = tpack_bf16(u0.z, u0.w);
                    w.z = tpack_bf16(u1.x, u1.y);
                    w.w = tpack_bf16(u1.z, u1.w);
                    bop[s][nb] = __builtin_bit_cast(tbf16x8, w);
                } else {
                    bop[s][nb] = *reinterpret_cast<const float4 *>(ua);
                }
            }
        __syncthreads();
    }

    for (long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const long tile = grp * 4 + wave;
        const long r0 = tile * 128;
        const bool live = tile < ntiles;
        tf32x4 acc[8][NBA];
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int nb = 0; nb < NBA; ++nb) acc[b][nb] = tf32x4{0.f, 0.f, 0.f, 0.f};
        if (live) {
            // per-lane base offsets of the lane's row groups
            long xb[2];
            bool rok[2];
            if constexpr (TR) {
                constexpr int G = (ST == 1) ? 8 : 4;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const long k = r0 + ((ST == 1) ? 8 * lm : 64 * h + 4 * lm);
                    rok[h] = k < rows;
                    const long kk = rok[h] ? k : 0;
                    const long q = kk / P, p = kk - q * P;
                    xb[h] = p + q * P * In;
                    (void)G;
                }
            }
            // chunks of CH steps: all loads of a chunk are issued before its MFMAs (steps past
            // In load clamped addresses and are masked to zero)
            constexpr int CH = 4;
#pragma unroll
            for (int c0 = 0; c0 < NSTEP; c0 += CH) {
                if (c0 >= nsteps) break;
                VT v[CH][8];
                uint32_t ok[CH];
#pragma unroll
                for (int cs_ = 0; cs_ < CH; ++cs_) {
                    const int s = c0 + cs_;
                    ok[cs_] = 0;
                    if constexpr (TR && ST == 1) {
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const long i = (long)s * KS + 8 * lg + j;
                            const bool in = rok[0] && i < In;
                            v[cs_][j] = ld_stream(reinterpret_cast<const VT *>(X + xb[0] + (in ? i : 0) * P));
                            ok[cs_] |= (uint32_t)in << j;
                        }
                    } else if constexpr (TR) {
#pragma unroll
                        for (int h = 0; h < 2; ++h)
#pragma unroll
                            for (int c = 0; c < 4; ++c) {
                                const long i = (long)s * KS + 4 * lg + c;
                                const bool in = rok[h] && i < In;
                                v[cs_][4 * h + c] = ld_stream(reinterpret_cast<const VT *>(X + xb[h] + (in ? i : 0) * P));
                                ok[cs_] |= (uint32_t)in << (4 * h + c);
                            }
                    } else {
                        const long i = (long)s * KS + ((ST == 1) ? 8 : 4) * lg;
#pragma unroll
                        for (int b = 0; b < 8; ++b) {
                            const long q = r0 + 16 * b + lm;
                            const bool in = q < rows && i < In;
                            v[cs_][b] = ld_stream(reinterpret_cast<const VT *>(X + (in ? i + q * In : 0)));
                            ok[cs_] |= (uint32_t)in << b;
                        }
                    }
                }
                __builtin_amdgcn_sched_barrier(0);  // all loads of the chunk before its MFMAs
#pragma unroll
                for (int cs_ = 0; cs_ < CH; ++cs_) {
                    const int s = c0 + cs_;
                    if constexpr (ST == 1) {
                        tbf16x8 op[8];
                        if constexpr (TR) {
                            uint4 w[8];
#pragma unroll
                            for (int j = 0; j < 8; ++j) w[j] = tmask(v[cs_][j], (ok[cs_] >> j) & 1u);
#pragma unroll
                            for (int t = 0; t < 8; ++t) {
                                uint4 o;
                                o.x = th16(w[0], t) | (th16(w[1], t) << 16);
                                o.y = th16(w[2], t) | (th16(w[3], t) << 16);
                                o.z = th16(w[4], t) | (th16(w[5], t) << 16);
                                o.w = th16(w[6], t) | (th16(w[7], t) << 16);
                                op[t] = __builtin_bit_cast(tbf16x8, o);
                            }
                        } else {
#pragma unroll
                            for (int b = 0; b < 8; ++b)
                                op[b] = __builtin_bit_cast(tbf16x8, tmask(v[cs_][b], (ok[cs_] >> b) & 1u));
                        }
#pragma unroll
                        for (int b = 0; b < 8; ++b)
#pragma unroll
                            for (int nb = 0; nb < NBA; ++nb)
                                acc[b][nb] =
                                    __builtin_amdgcn_mfma_f32_16x16x32_bf16(op[b], bop[s][nb], acc[b][nb], 0, 0, 0);
                    } else {
                        float4 w[8];
#pragma unroll
                        for (int j = 0; j < 8; ++j) w[j] = tmask(v[cs_][j], (ok[cs_] >> j) & 1u);
#pragma unroll
                        for (int c = 0; c < 4; ++c)
#pragma unroll
                            for (int b = 0; b < 8; ++b) {
                                const float a = TR ? tf4(w[4 * (b >> 2) + c], b & 3) : tf4(w[b], c);
#pragma unroll
                                for (int nb = 0; nb < NBA; ++nb)
                                    acc[b][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, tf4(bop[s][nb], c),
                                                                                      acc[b][nb], 0, 0, 0);
                            }
                    }
                }
            }
        }
        // stage 32 columns at a time in LDS (this wave's slice), then contiguous stores
#pragma unroll
        for (int hh = 0; hh < (NBA + 1) / 2; ++hh) {
#pragma unroll
            for (int nb2 = 0; nb2 < 2; ++nb2) {
                const int nb = 2 * hh + nb2;
                if (nb < NBA) {
#pragma unroll
                    for (int b = 0; b < 8; ++b)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            cw[(16 * nb2 + lm) * kTtmCsLd + ttm_row<ST, TR>(b, 4 * lg + v)] = acc[b][nb][v];
                }
            }
            __syncthreads();
            if (live) {
                const long acols = min(32L, Ra - 32L * hh);
                if constexpr (TR) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const long k = r0 + lane + 64 * h;
                        if (k < rows) {
                            const long q = k / P, p = k - q * P;
                            float *yb = Y + p + q * P * Ra + (32L * hh) * P;
                            for (long a = 0; a < acols; ++a) yb[a * P] = cw[a * kTtmCsLd + lane + 64 * h];
                        }
                    }
                } else {
                    // rows r0 .. r0+127 of Y (Ra floats each) are contiguous
                    if (acols == 32) {
                        for (int e = lane; e < 128 * 32; e += 64) {
                            const int r = e >> 5, a = e & 31;
                            if (r0 + r < rows) Y[(r0 + r) * Ra + 32L * hh + a] = cw[a * kTtmCsLd + r];
                        }
                    } else {
                        for (int r = 0; r < 128; ++r)
                            if (lane < acols && r0 + r < rows) Y[(r0 + r) * Ra + 32L * hh + lane] = cw[lane * kTtmCsLd + r];
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Y = X x_n U^T on MFMA when the shape is covered (In <= 128, Ra <= 64, 16-byte vectors
// along the contiguous index); returns 1 if launched, 0 if not covered, -1 on error.
int ttm_mfma_launch(const void *X, int xtype, long P, long In, long Q, const float *U, long lda, long ldi, long Ra,
                    float *Y, hipStream_t s) {
    if (xtype != TDX_F32 && xtype != TDX_BF16) return 0;
    const int st = (xtype == TDX_BF16) ? 1 : 0;
    if (In < 1 || In > kTtmMaxIn || Ra < 1 || Ra > 64 || P < 1 || Q < 1) return 0;
    if ((uintptr_t)X & 15) return 0;
    const long vec = st ? 8 : 4;
    const bool tr = P > 1;
    if (tr && (P % vec)) return 0;
    if (!tr && (In % vec)) return 0;
    const int nba = (Ra <= 16) ? 1 : (Ra <= 32) ? 2 : 4;
    if (!st && nba > 2) return 0;  // FP32 B operands of 64 columns do not fit the register budget
    const long rows = tr ? P * Q : Q;
    const long groups = ceil_div(ceil_div(rows, 128L), 4L);
    const long grid = groups < 2048 ? groups : 2048;
    const dim3 g((unsigned)grid), b(256);
#define TDX_TTM_LAUNCH(ST_, TR_, NBA_) \
    k_ttm_mfma<ST_, TR_, NBA_><<<g, b, 0, s>>>(X, P, In, Q, U, lda, ldi, Ra, Y)
    if (st) {
        if (tr) {
            if (nba == 1) TDX_TTM_LAUNCH(1, true, 1);
            else if (nba == 2) TDX_TTM_LAUNCH(1, true, 2);
            else TDX_TTM_LAUNCH(1, true, 4);
        } else {
            if (nba == 1) TDX_TTM_LAUNCH(1, false, 1);
            else if (nba == 2) TDX_TTM_LAUNCH(1, false, 2);
            else TDX_TTM_LAUNCH(1, false, 4);
        }
    } else {
        if (tr) {
            if (nba == 1) TDX_TTM_LAUNCH(0, true, 1);
            else TDX_TTM_LAUNCH(0, true, 2);
        } else {
            if (nba == 1) TDX_TTM_LAUNCH(0, false, 1);
            else TDX_TTM_LAUNCH(0, false, 2);
        }
    }
#undef TDX_TTM_LAUNCH
    if (hipGetLastError() != hipSuccess) return -1;
    return 1;
}

}  // namespace tdx
