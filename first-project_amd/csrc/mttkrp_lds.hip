// mttkrp_lds.hip -- fused MTTKRP for an output mode n > 0 (the contraction runs along the
// contiguous mode 0), with the tensor tile and the Khatri-Rao block staged in LDS.
//
// M[i_n, r] = sum_{i_0, slow} X[i_0, .., i_n, ..] * P_0[i_0, r] * w_slow[r],
// w_slow[r] = prod over the other contracted modes of their factor rows.
//
// Work of a workgroup: 128 consecutive output rows and a range of slow indices; for every
// chunk of KB consecutive i_0 (KB = 128 BF16/FP8, 64 FP32) it walks its slow indices:
//   * the (128 x KB) tensor tile is read with 16-byte loads, 16 (FP8: 8) lanes per row, so
//     each row contributes whole 128-byte lines (256 bytes, FP8 128) per step -- output rows
//     are far apart in memory (the stride of mode n), and half-line accesses spread over
//     time were re-fetched;
//   * the KB x Rp Khatri-Rao block  P_0[chunk, :] * w_slow  is formed once per step in LDS
//     for all four waves (each thread keeps its 32 panel values of the chunk in registers
//     for the whole slow loop);
//   * each wave multiplies its 32 rows by the block (v_mfma_f32_16x16x32_bf16; FP32
//     storage v_mfma_f32_16x16x4_f32), accumulating in FP32;
//   * the tile of the next slow index is loaded into registers while the current one is
//     multiplied (one step in flight per workgroup, two workgroups per CU).
// Waves own disjoint rows, so each writes its rows of the workgroup's FP32 partial directly;
// the partial layout [split][row][Rp] is the one of the register-tile kernel.
#include "mttkrp.hpp"

namespace tdx {

typedef float lf32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 lbf16x8 __attribute__((ext_vector_type(8)));
typedef float lf32x2 __attribute__((ext_vector_type(2)));

struct LArgs {
    const void *X;
    const float *panel[4];
    float *partial;
    const int *skip;
    long stride[4];
    long dims[4];
    long In, If, S, sper, rows_pad;  // S slow indices in total, sper per split
    int order, mode, nslow, Rp, rowtiles;
    int slow[2];
};

__device__ __forceinline__ uint32_t lpack_bf16(float a, float b) {
    auto rne = [](float f) -> uint32_t {
        const uint32_t u = __float_as_uint(f);
        if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
        return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
    };
    return rne(a) | (rne(b) << 16);
}

// 4 FP8 E4M3 codes -> 4 BF16 (exact)
__device__ __forceinline__ uint2 lfp8x4_to_bf16(uint32_t q) {
    const lf32x2 a = __builtin_amdgcn_cvt_pk_f32_fp8((int)q, false);
    const lf32x2 b = __builtin_amdgcn_cvt_pk_f32_fp8((int)q, true);
    return uint2{(__float_as_uint(a.x) >> 16) | (__float_as_uint(a.y) & 0xffff0000u),
                 (__float_as_uint(b.x) >> 16) | (__float_as_uint(b.y) & 0xffff0000u)};
}

template <int ST> struct LTile;
// FP32: KB = 64, a row is 256 bytes = 16 lanes x 16 bytes; LDS keeps FP32
template <> struct LTile<0> { static constexpr int KB = 64, LPR = 16, PAD = 4; typedef float E; };
// BF16: KB = 128, 256 bytes per row = 16 lanes x 16 bytes; LDS keeps BF16
template <> struct LTile<1> { static constexpr int KB = 128, LPR = 16, PAD = 8; typedef uint16_t E; };
// FP8: KB = 128, 128 bytes per row = 8 lanes x 16 bytes; widened to BF16 in LDS
template <> struct LTile<2> { static constexpr int KB = 128, LPR = 8, PAD = 8; typedef uint16_t E; };

template <int ST, int NB>
__global__ __launch_bounds__(256, 2) void k_mttkrp_lds(LArgs a) {
    constexpr int KB = LTile<ST>::KB;
    constexpr int LPR = LTile<ST>::LPR;          // lanes per row in the global loads
    constexpr int RPP = 256 / LPR;                // rows per load pass
    constexpr int NPASS = 128 / RPP;              // 16-byte loads per thread per step
    constexpr int XLD = KB + LTile<ST>::PAD;      // LDS row pitch (elements)
    typedef typename LTile<ST>::E LE;
    constexpr int Rp = 16 * NB;
    __shared__ __attribute__((aligned(16))) LE xs[128 * XLD];  // tensor tile [row][k]
    __shared__ __attribute__((aligned(16))) LE bs[Rp * XLD];   // KR block [col][k]
    if (a.skip && *a.skip) return;

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lm = lane & 15, lg = lane >> 4;
    const int rowtile = blockIdx.x % a.rowtiles;
    const long split = blockIdx.x / a.rowtiles;
    const long row0 = (long)rowtile * 128;
    const long sl0 = split * a.sper;
    const long sl1 = min(sl0 + a.sper, a.S);
    const long nfc = ceil_div(a.If, (long)KB);
    const long prows = ceil_div(a.If, 64L) * 64;  // panel rows (zero padded to a multiple of 64)
    const long sn = a.stride[a.mode];

    // global-load geometry: this thread loads 16 bytes of rows  tid / LPR + RPP * p
    const int lrow = tid / LPR, lcol = (tid % LPR) * ((ST == 0) ? 4 : (ST == 1) ? 8 : 16);
    long roff[NPASS];
    bool rok[NPASS];
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
        const long r = row0 + lrow + RPP * p;
        rok[p] = r < a.In;
        roff[p] = (rok[p] ? r : 0) * sn;
    }
    // KR-block geometry: this thread forms column kr_c, elements kr_k0 .. kr_k0 + KB/4 - 1
    // (Rp = 64; smaller Rp leaves threads idle)
    const int kr_c = tid >> 2, kr_k0 = (tid & 3) * (KB / 4);
    const bool kr_on = kr_c < Rp;

    lf32x4 acc[2][NB];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[rb][nb] = lf32x4{0.f, 0.f, 0.f, 0.f};

    // slow index -> tensor offset and this thread's weight w_slow[kr_c]
    // (at most two slow modes; both factor rows are loaded before either is used)
    const int sm0 = a.slow[0], sm1 = (a.nslow > 1) ? a.slow[1] : a.slow[0];
    const long sd0 = a.dims[sm0];
    // (the two factor values are multiplied only when the block is formed, so the loads
    // are not waited for before the MFMAs of the current step)
    auto slow_of = [&](long sl, long &off, float &w0, float &w1) {
        const long i0 = sl % sd0, i1 = (a.nslow > 1) ? sl / sd0 : 0;
        off = i0 * a.stride[sm0] + ((a.nslow > 1) ? i1 * a.stride[sm1] : 0);
        const int c = kr_on ? kr_c : 0;
        w0 = a.panel[sm0][((i0 >> 3) * Rp + c) * 8 + (i0 & 7)];
        w1 = a.panel[sm1][((i1 >> 3) * Rp + c) * 8 + (i1 & 7)];
    };

    typedef uint4 VT;
    VT xr[NPASS];
    uint32_t xok = 0;
    float wa = 0.0f, wb = 0.0f;
    auto load_tile = [&](long fc, long sl) {
        long off;
        slow_of(sl, off, wa, wb);
        const long f = fc * KB + lcol;
        const bool fok = f < a.If;
        xok = 0;
#pragma unroll
        for (int p = 0; p < NPASS; ++p) {
            const bool in = rok[p] && fok;
            const LE *src;
            if constexpr (ST == 0) src = reinterpret_cast<const LE *>(reinterpret_cast<const float *>(a.X) + off + roff[p] + (fok ? f : 0));
            else if constexpr (ST == 1) src = reinterpret_cast<const LE *>(reinterpret_cast<const uint16_t *>(a.X) + off + roff[p] + (fok ? f : 0));
            else src = reinterpret_cast<const LE *>(reinterpret_cast<const uint8_t *>(a.X) + off + roff[p] + (fok ? f : 0));
            xr[p] = *reinterpret_cast<const VT *>(src);
            xok |= (uint32_t)in << p;
        }
    };

    const long nsl = sl1 - sl0;
    for (long fc = 0; fc < nfc && nsl > 0; ++fc) {
        // this thread's panel values of the chunk, kept for the whole slow loop
        float pc[KB / 4];
#pragma unroll
        for (int g = 0; g < KB / 32; ++g) {
            // rows past the panel's padding (a multiple of 64) are zero
            const long i = fc * KB + kr_k0 + 8 * g;
            const bool pin = kr_on && i < prows;
            const float *pp = a.panel[0] + (((pin ? i : 0) >> 3) * Rp + (kr_on ? kr_c : 0)) * 8;
            float4 u0 = *reinterpret_cast<const float4 *>(pp);
            float4 u1 = *reinterpret_cast<const float4 *>(pp + 4);
            if (!pin) u0 = u1 = float4{0.f, 0.f, 0.f, 0.f};
            pc[8 * g + 0] = u0.x; pc[8 * g + 1] = u0.y; pc[8 * g + 2] = u0.z; pc[8 * g + 3] = u0.w;
            pc[8 * g + 4] = u1.x; pc[8 * g + 5] = u1.y; pc[8 * g + 6] = u1.z; pc[8 * g + 7] = u1.w;
        }
        load_tile(fc, sl0);
        for (long s = 0; s < nsl; ++s) {
            __syncthreads();  // every wave is done with the previous tile and block
            // tensor tile -> LDS (masked lanes store zeros)
#pragma unroll
            for (int p = 0; p < NPASS; ++p) {
                const bool in = (xok >> p) & 1u;
                const uint4 v = in ? xr[p] : uint4{0u, 0u, 0u, 0u};
                LE *dst = xs + (lrow + RPP * p) * XLD + lcol;
                if constexpr (ST == 2) {
                    const uint2 b0 = lfp8x4_to_bf16(v.x), b1 = lfp8x4_to_bf16(v.y);
                    const uint2 b2 = lfp8x4_to_bf16(v.z), b3 = lfp8x4_to_bf16(v.w);
                    *reinterpret_cast<uint4 *>(dst) = uint4{b0.x, b0.y, b1.x, b1.y};
                    *reinterpret_cast<uint4 *>(dst + 8) = uint4{b2.x, b2.y, b3.x, b3.y};
                } else {
                    *reinterpret_cast<uint4 *>(dst) = v;
                }
            }
            // Khatri-Rao block column kr_c: P_0[chunk, kr_c] * w
            if (kr_on) {
                const float w = (a.nslow > 1) ? wa * wb : wa;
                LE *dst = bs + kr_c * XLD + kr_k0;
                if constexpr (ST == 0) {
#pragma unroll
                    for (int q = 0; q < KB / 16; ++q)
                        *reinterpret_cast<float4 *>(dst + 4 * q) =
                            float4{pc[4 * q] * w, pc[4 * q + 1] * w, pc[4 * q + 2] * w, pc[4 * q + 3] * w};
                } else {
#pragma unroll
                    for (int q = 0; q < KB / 32; ++q)
                        *reinterpret_cast<uint4 *>(dst + 8 * q) =
                            uint4{lpack_bf16(pc[8 * q] * w, pc[8 * q + 1] * w), lpack_bf16(pc[8 * q + 2] * w, pc[8 * q + 3] * w),
                                  lpack_bf16(pc[8 * q + 4] * w, pc[8 * q + 5] * w), lpack_bf16(pc[8 * q + 6] * w, pc[8 * q + 7] * w)};
                }
            }
            __syncthreads();
            // next tile in flight while this one is multiplied (the last step of a chunk
            // re-reads its own tile: never consumed, cache hit)
            load_tile(fc, sl0 + ((s + 1 < nsl) ? s + 1 : s));
            __builtin_amdgcn_sched_barrier(0);  // the next tile's loads stay ahead of the MFMAs
            // MFMAs: this wave's rows 32 wave + 16 rb + (lane % 16)
            if constexpr (ST == 0) {
#pragma unroll
                for (int q = 0; q < KB / 16; ++q) {
                    float4 av[2], bv[NB];
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
                        av[rb] = *reinterpret_cast<const float4 *>(xs + (32 * wave + 16 * rb + lm) * XLD + 16 * q + 4 * lg);
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        bv[nb] = *reinterpret_cast<const float4 *>(bs + (16 * nb + lm) * XLD + 16 * q + 4 * lg);
#pragma unroll
                    for (int c = 0; c < 4; ++c)
#pragma unroll
                        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                            for (int nb = 0; nb < NB; ++nb) {
                                const float x = c == 0 ? av[rb].x : c == 1 ? av[rb].y : c == 2 ? av[rb].z : av[rb].w;
                                const float y = c == 0 ? bv[nb].x : c == 1 ? bv[nb].y : c == 2 ? bv[nb].z : bv[nb].w;
                                acc[rb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, acc[rb][nb], 0, 0, 0);
                            }
                }
            } else {
#pragma unroll
                for (int q = 0; q < KB / 32; ++q) {
                    lbf16x8 av[2], bv[NB];
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
                        av[rb] = __builtin_bit_cast(
                            lbf16x8, *reinterpret_cast<const uint4 *>(xs + (32 * wave + 16 * rb + lm) * XLD + 32 * q + 8 * lg));
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        bv[nb] = __builtin_bit_cast(
                            lbf16x8, *reinterpret_cast<const uint4 *>(bs + (16 * nb + lm) * XLD + 32 * q + 8 * lg));
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb)
                            acc[rb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[rb], bv[nb], acc[rb][nb], 0, 0, 0);
                }
            }
        }
    }
    // this wave's 32 rows of the partial: lane holds rows 4 lg + v of each 16-row block
    float *dst = a.partial + (split * a.rows_pad + row0 + 32 * wave) * Rp;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int v = 0; v < 4; ++v) dst[(16 * rb + 4 * lg + v) * Rp + 16 * nb + lm] = acc[rb][nb][v];
}

// covered when the output mode is not 0, the rank is <= 64 and at most two other modes are
// contracted besides mode 0 (orders 3 and 4)
bool mttkrp_lds_covers(int order, int mode, long rank) { return mode > 0 && rank <= 64 && order >= 3 && order <= 4; }

int mttkrp_lds_launch(const MttkrpPlan &p, const void *X, const float *const *panels, float *partial,
                      hipStream_t s, const int *skip) {
    LArgs a{};
    a.X = X;
    for (int m = 0; m < 4; ++m) a.panel[m] = (m < p.order) ? panels[m] : nullptr;
    a.partial = partial;
    a.skip = skip;
    long st = 1;
    for (int m = 0; m < 4; ++m) {
        a.dims[m] = (m < p.order) ? p.dims[m] : 1;
        a.stride[m] = st;
        st *= a.dims[m];
    }
    a.order = p.order;
    a.mode = p.mode;
    a.In = p.dims[p.mode];
    a.If = p.dims[0];
    a.Rp = p.Rp;
    a.rowtiles = p.rowtiles;
    a.rows_pad = p.rows_pad;
    a.S = p.ksteps;             // slow indices in total
    a.sper = p.ksteps_per_split;
    a.nslow = 0;
    for (int m = 1; m < p.order; ++m)
        if (m != p.mode) a.slow[a.nslow++] = m;
    const int grid = p.rowtiles * p.nsplit;
    const int stt = (p.dtype == TDX_F32) ? 0 : (p.dtype == TDX_BF16) ? 1 : 2;
#define TDX_LDS_LAUNCH(ST_)                                                       \
    switch (p.Rp / 16) {                                                          \
        case 1: k_mttkrp_lds<ST_, 1><<<grid, 256, 0, s>>>(a); break;              \
        case 2: k_mttkrp_lds<ST_, 2><<<grid, 256, 0, s>>>(a); break;              \
        case 3: k_mttkrp_lds<ST_, 3><<<grid, 256, 0, s>>>(a); break;              \
        default: k_mttkrp_lds<ST_, 4><<<grid, 256, 0, s>>>(a); break;             \
    }
    if (stt == 0) { TDX_LDS_LAUNCH(0) }
    else if (stt == 1) { TDX_LDS_LAUNCH(1) }
    else { TDX_LDS_LAUNCH(2) }
#undef TDX_LDS_LAUNCH
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

}  // namespace tdx
