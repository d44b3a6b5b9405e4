// mttkrp_lds.hip -- fused MTTKRP with the tensor tile and the Khatri-Rao block staged in
// LDS.  For an output mode n > 0 the contraction's fast index f is the contiguous mode 0;
// for n = 0 (TR) it is mode 1 and the tile is transposed on its way into LDS.
//
// M[i_n, r] = sum_{i_f, slow} X[...] * P_f[i_f, r] * w_slow[r],
// w_slow[r] = prod over the other contracted modes of their factor rows.
//
// Work of a workgroup: 128 consecutive output rows and a range of slow indices; for every
// chunk of KB consecutive i_f (KB = 128 BF16/FP8, 64 FP32) it walks its slow indices:
//   * n > 0: the (128 x KB) tensor tile is read with 16-byte loads, 16 (FP8: 8) lanes per
//     row, so each row contributes whole 128-byte lines (256 bytes, FP8 128) per step --
//     output rows are far apart in memory (the stride of mode n), and half-line accesses
//     spread over time were re-fetched;
//   * n = 0 (TR): output rows are contiguous; a thread loads 8 consecutive i_1 of 8 rows
//     (4 rows for FP32) with 16-byte (FP8 8-byte) loads, 16 threads covering 256 contiguous
//     bytes per i_1, and transposes the block in registers so LDS holds [row][k] as above.
//     Row RL g + t of the tile (row group g) sits in LDS row RG t + g, so the 16-byte stores
//     of a wave go to consecutive LDS rows (no bank conflicts); the MFMA rows follow the LDS
//     order and are mapped back to tile rows when the partial is written;
//   * n = 0, FP8 storage: the tile stays FP8 in LDS in its global order, [i_1][128 rows]
//     (16 KB instead of 32 KB widened; 16-byte loads, 8 lanes per whole 128-byte line, stored
//     as they arrive), and the MFMA A operand comes from ds_read_b64_tr_b8 (8 rows x 8 i_1
//     bytes per 16-lane group, transposed) widened to BF16 in registers -- no register
//     transpose, half the tile's LDS bytes;
//   * the KB x Rp Khatri-Rao block  P_0[chunk, :] * w_slow  is formed once per step in LDS
//     for all four waves (each thread keeps its 32 panel values of the chunk in registers
//     for the whole slow loop);
//   * each wave multiplies its 32 rows by the block (v_mfma_f32_16x16x32_bf16; FP32
//     storage v_mfma_f32_16x16x4_f32), accumulating in FP32;
//   * the tiles of the next two slow indices are in flight in two register buffers while the
//     current one is multiplied (two steps in flight per workgroup, two workgroups per CU);
//   * a workgroup's work is one i_0 chunk (rarely a few) times a range of slow indices, so
//     the pipeline runs unbroken over the whole range.
// Waves own disjoint rows, so each writes its rows of the workgroup's FP32 partial directly;
// the partial layout [split][row][Rp] is the one of the register-tile kernel.
#include "mttkrp.hpp"
#include <type_traits>

namespace tdx {

typedef float lf32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 lbf16x8 __attribute__((ext_vector_type(8)));
typedef float lf32x2 __attribute__((ext_vector_type(2)));

struct LArgs {
    const void *X;
    const float *panel[4];
    const float *fpanel;     // panel of the fast mode (mode 0, or mode 1 when n = 0)
    const float *row_scale;  // FP8 storage, n = 0: per-row scales applied to the output
    float *partial;
    const int *skip;
    long stride[4];
    long dims[4];
    long In, If, S, sper, rows_pad;  // If: fast extent; S slow indices in total, sper per split
    int order, mode, nslow, Rp, rowtiles;
    int cps, nss, ncg;                // chunks per split, slow ranges per chunk group, chunk groups
    int slow[2];
};

typedef __bf16 lbf16x2 __attribute__((ext_vector_type(2)));
// two FP32 -> packed BF16, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t lpack_bf16(float a, float b) {
    lbf16x2 v;
    v.x = (__bf16)a;
    v.y = (__bf16)b;
    return __builtin_bit_cast(uint32_t, v);
}

// 4 FP8 E4M3 codes -> 4 BF16 (exact: v_cvt_scalef32_pk_bf16_fp8 with scale 1, two codes
// per instruction, bit-identical to widening through FP32 for all 256 codes)
__device__ __forceinline__ uint2 lfp8x4_to_bf16(uint32_t q) {
    return uint2{__builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(q, 1.0f, false)),
                 __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(q, 1.0f, true))};
}

template <int ST> struct LTile;
// FP32: KB = 64, a row is 256 bytes = 16 lanes x 16 bytes; LDS keeps FP32
template <> struct LTile<0> { static constexpr int KB = 64, LPR = 16, PAD = 4; typedef float E; };
// BF16: KB = 128, 256 bytes per row = 16 lanes x 16 bytes; LDS keeps BF16
template <> struct LTile<1> { static constexpr int KB = 128, LPR = 16, PAD = 8; typedef uint16_t E; };
// FP8: KB = 128, 128 bytes per row = 8 lanes x 16 bytes; widened to BF16 in LDS
template <> struct LTile<2> { static constexpr int KB = 128, LPR = 8, PAD = 8; typedef uint16_t E; };

// 8 FP8 E4M3 codes -> 8 BF16
__device__ __forceinline__ uint4 lfp8x8_to_bf16(uint2 q) {
    const uint2 lo = lfp8x4_to_bf16(q.x), hi = lfp8x4_to_bf16(q.y);
    return uint4{lo.x, lo.y, hi.x, hi.y};
}
// 16-bit element t of an 8-element vector held in a uint4
__device__ __forceinline__ uint32_t lhalf(const uint4 &v, int t) {
    const uint32_t w = (t >> 1) == 0 ? v.x : (t >> 1) == 1 ? v.y : (t >> 1) == 2 ? v.z : v.w;
    return (t & 1) ? (w >> 16) : (w & 0xffffu);
}
__device__ __forceinline__ uint32_t lword(const uint4 &v, int t) {
    return t == 0 ? v.x : t == 1 ? v.y : t == 2 ? v.z : v.w;
}

// TR (n = 0) load geometry: a load covers RL rows of one i_1; RG row groups x KG k groups
template <int ST> struct LTr {
    static constexpr int RL = (ST == 0) ? 4 : 8;
    static constexpr int RG = 128 / RL;
    static constexpr int KG = 256 / RG;
    typedef typename std::conditional<ST == 2, uint2, uint4>::type V;
};

// SPLIT (BF16 storage only): the Khatri-Rao block is kept as hi + lo BF16 parts
// (hi = bf16(v), lo = bf16(v - hi)) and every tile is multiplied by both, so the block is
// used to ~2^-17 relative instead of 2^-9: the MTTKRP of the BF16-rounded tensor with
// near-FP32 factors (twice the MFMAs and block reads; the tensor traffic is unchanged).
template <int ST, int NB, bool TR, bool SPLIT = false>
__global__ __launch_bounds__(256, 2) void k_mttkrp_lds(LArgs a) {
    static_assert(!SPLIT || ST == 1, "the split Khatri-Rao block is for BF16 storage");
    constexpr bool F8T = (ST == 2) && TR;       // FP8 mode 0: FP8 tile [k][row] in LDS
    constexpr int KB = LTile<ST>::KB;
    constexpr int LPR = LTile<ST>::LPR;          // lanes per row in the global loads (n > 0)
    constexpr int RPP = 256 / LPR;                // rows per load pass
    constexpr int NPASS = F8T ? 4 : TR ? 8 : 128 / RPP;  // loads per thread per step
    static_assert(!TR || 8 * LTr<ST>::KG == KB, "TR tile covers KB fast indices");
    constexpr int XLD = KB + LTile<ST>::PAD;      // LDS row pitch (elements)
    typedef typename LTile<ST>::E LE;
    constexpr int Rp = 16 * NB;
    // tensor tile: [row][k] (F8T: 128 k-rows of 128 FP8 bytes, 16-byte granules swizzled)
    __shared__ __attribute__((aligned(16))) LE xs[F8T ? 128 * 128 / 2 : 128 * XLD];
    __shared__ __attribute__((aligned(16))) LE bs[(SPLIT ? 2 : 1) * Rp * XLD];  // KR block [col][k] (+ lo part)
    if (a.skip && *a.skip) return;

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lm = lane & 15, lg = lane >> 4;
    // Workgroups that run side by side should read whole rows of the tensor between them
    // (DRAM page locality): for TR adjacent workgroups take adjacent row tiles (neighbouring
    // 256-byte pieces of the same i_1 rows), otherwise adjacent chunk groups (the pieces of
    // the same output rows).
    int rowtile;
    long cgroup, srange;
    if constexpr (TR) {
        rowtile = blockIdx.x % a.rowtiles;
        const long split = blockIdx.x / a.rowtiles;
        cgroup = split / a.nss;
        srange = split % a.nss;
    } else {
        cgroup = blockIdx.x % a.ncg;
        const long rest = blockIdx.x / a.ncg;
        rowtile = (int)(rest % a.rowtiles);
        srange = rest / a.rowtiles;
    }
    const long split = cgroup * a.nss + srange;
    const long row0 = (long)rowtile * 128;
    const long nfc = ceil_div(a.If, (long)KB);
    const long fc0 = cgroup * a.cps, fc1 = min(fc0 + a.cps, nfc);
    const long sl0 = srange * a.sper;
    const long sl1 = min(sl0 + a.sper, a.S);
    const long prows = ceil_div(a.If, 64L) * 64;  // panel rows (zero padded to a multiple of 64)
    const long sn = a.stride[a.mode];
    const long sf = TR ? a.stride[1] : 1;          // stride of the fast index

    // global-load geometry.  n > 0: this thread loads 16 bytes of rows  tid / LPR + RPP * p
    // at fast offset lcol.  TR: rows RL * (tid % RG) .. + RL - 1 at fast offsets
    // 8 * (tid / RG) + p (the row group is whole: the planner requires In % RL == 0).
    // F8T: this thread loads rows 16 (tid % 8) .. + 15 (16 bytes) of fast index tid / 8 + 32 p
    // (8 lanes read one whole 128-byte line per fast index; the planner requires In % 16 == 0).
    const int lrow = F8T ? 16 * (tid % 8) : TR ? LTr<ST>::RL * (tid % LTr<ST>::RG) : tid / LPR;
    const int lcol = F8T ? tid / 8 : TR ? 8 * (tid / LTr<ST>::RG) : (tid % LPR) * ((ST == 0) ? 4 : (ST == 1) ? 8 : 16);
    long roff[NPASS];
    bool rok[NPASS];
#pragma unroll
    for (int p = 0; p < NPASS; ++p) {
        const long r = TR ? row0 + lrow : row0 + lrow + RPP * p;
        rok[p] = r < a.In;
        roff[p] = (rok[p] ? r : 0) * (TR ? 1 : sn);
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

    // The slow-index frontier: the next slow index to load as (i0, i1) digits, its tensor
    // offset, and how many indices of the range are still ahead (at the last one the
    // frontier stays put, so the pipeline's surplus loads repeat it).  Each load advances
    // it by one, so no division is needed inside the loop.  (At most two slow modes.)
    const int sm0 = a.slow[0], sm1 = (a.nslow > 1) ? a.slow[1] : a.slow[0];
    const long sd0 = a.dims[sm0];
    const long st0 = a.stride[sm0], st1 = (a.nslow > 1) ? a.stride[sm1] : 0;
    long f_i0 = 0, f_i1 = 0, f_off = 0, f_left = 0;
    auto frontier_at = [&](long sl, long count) {
        f_i0 = sl % sd0;
        f_i1 = (a.nslow > 1) ? sl / sd0 : 0;
        f_off = f_i0 * st0 + f_i1 * st1;
        f_left = count;
    };
    const int wc = kr_on ? kr_c : 0;

    typedef typename std::conditional<TR && !F8T, typename LTr<ST>::V, uint4>::type VT;
    // one slow step's tensor loads (masked lanes recorded in ok) and its slow-factor values
    struct TileBuf {
        VT x[NPASS];
        uint32_t ok;
        float wa, wb;
    };
    // loads the frontier's tile of chunk fc into b, then advances the frontier
    auto load_tile = [&](TileBuf &b, long fc) {
        const long off = f_off;
        // this thread's weights (multiplied only when the block is formed, so the loads are
        // not waited for before the MFMAs of the current step)
        b.wa = a.panel[sm0][((f_i0 >> 3) * Rp + wc) * 8 + (f_i0 & 7)];
        b.wb = a.panel[sm1][((f_i1 >> 3) * Rp + wc) * 8 + (f_i1 & 7)];
        if (f_left > 1) {
            --f_left;
            ++f_i0;
            f_off += st0;
            if (f_i0 == sd0) {
                f_i0 = 0;
                ++f_i1;
                f_off += st1 - sd0 * st0;
            }
        }
        b.ok = 0;
#pragma unroll
        for (int p = 0; p < NPASS; ++p) {
            // fast index of this load; out-of-range loads read the tile origin instead
            const long f = fc * KB + lcol + (F8T ? 32 * p : TR ? p : 0);
            const bool fok = f < a.If;
            const bool in = rok[p] && fok;
            const long e = off + roff[p] + (fok ? f : 0) * sf;
            const LE *src;
            if constexpr (ST == 0) src = reinterpret_cast<const LE *>(reinterpret_cast<const float *>(a.X) + e);
            else if constexpr (ST == 1) src = reinterpret_cast<const LE *>(reinterpret_cast<const uint16_t *>(a.X) + e);
            else src = reinterpret_cast<const LE *>(reinterpret_cast<const uint8_t *>(a.X) + e);
            b.x[p] = ld_stream(reinterpret_cast<const VT *>(src));
            b.ok |= (uint32_t)in << p;
        }
    };

    float pc[KB / 4];  // this thread's panel values of the current chunk
    // tensor tile -> LDS; MASKED stores zeros for out-of-range loads (needed only when the
    // tile crosses the edge of the rows or of the fast extent)
    auto write_tile = [&](const TileBuf &b, auto masked_tag) {
        constexpr bool MASKED = decltype(masked_tag)::value;
        if constexpr (F8T) {
            // fast index k = lcol + 32 p, 16-byte granule g = lrow / 16 of its 128-byte k-row,
            // stored at granule g ^ ((k >> 1) & 7): the 16 lanes of a write cover two whole
            // k-rows, and the 8 k-rows a transposing read takes (same g) hit 8 distinct
            // granule positions x two row parities -- no bank conflicts either way
            unsigned char *x8 = reinterpret_cast<unsigned char *>(xs);
#pragma unroll
            for (int p = 0; p < NPASS; ++p) {
                const int k = lcol + 32 * p;
                const uint4 v = (!MASKED || ((b.ok >> p) & 1u)) ? b.x[p] : uint4{0u, 0u, 0u, 0u};
                *reinterpret_cast<uint4 *>(x8 + k * 128 + 16 * ((lrow >> 4) ^ ((k >> 1) & 7))) = v;
            }
        } else if constexpr (TR) {
            if constexpr (ST == 0) {
                // 8 loads of 4 rows -> tile rows lrow + t, 8 fast indices each (two 16-byte stores)
                uint4 v[8];
#pragma unroll
                for (int p = 0; p < 8; ++p) v[p] = (!MASKED || ((b.ok >> p) & 1u)) ? b.x[p] : uint4{0u, 0u, 0u, 0u};
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    LE *dst = xs + (LTr<ST>::RG * t + tid % LTr<ST>::RG) * XLD + lcol;
                    *reinterpret_cast<uint4 *>(dst) = uint4{lword(v[0], t), lword(v[1], t), lword(v[2], t), lword(v[3], t)};
                    *reinterpret_cast<uint4 *>(dst + 4) = uint4{lword(v[4], t), lword(v[5], t), lword(v[6], t), lword(v[7], t)};
                }
            } else {
                // 8 loads of 8 rows (BF16, FP8 widened) -> 8x8 transpose of 16-bit elements
                uint4 v[8];
#pragma unroll
                for (int p = 0; p < 8; ++p) {
                    uint4 u;
                    if constexpr (ST == 2) u = lfp8x8_to_bf16(b.x[p]);
                    else u = b.x[p];
                    v[p] = (!MASKED || ((b.ok >> p) & 1u)) ? u : uint4{0u, 0u, 0u, 0u};
                }
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    uint4 o;
                    o.x = lhalf(v[0], t) | (lhalf(v[1], t) << 16);
                    o.y = lhalf(v[2], t) | (lhalf(v[3], t) << 16);
                    o.z = lhalf(v[4], t) | (lhalf(v[5], t) << 16);
                    o.w = lhalf(v[6], t) | (lhalf(v[7], t) << 16);
                    *reinterpret_cast<uint4 *>(xs + (LTr<ST>::RG * t + tid % LTr<ST>::RG) * XLD + lcol) = o;
                }
            }
        } else {
#pragma unroll
            for (int p = 0; p < NPASS; ++p) {
                const bool in = !MASKED || ((b.ok >> p) & 1u);
                const uint4 v = in ? b.x[p] : uint4{0u, 0u, 0u, 0u};
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
        }
    };
    // One step: buffer b (the tile of the current slow index) -> LDS, its KR block, then b is
    // refilled from the frontier while the MFMAs run (skipped when !valid: the padding step
    // of an odd-length range).  full: the tile lies inside the rows and the fast extent.
    auto step = [&](TileBuf &b, long fc, bool full, bool valid) {
        __syncthreads();  // every wave is done with the previous tile and block
        if (full) write_tile(b, std::false_type{});
        else write_tile(b, std::true_type{});
        // Khatri-Rao block column kr_c: P_0[chunk, kr_c] * w
        if (kr_on) {
            const float w = (a.nslow > 1) ? b.wa * b.wb : b.wa;
            LE *dst = bs + kr_c * XLD + kr_k0;
            if constexpr (ST == 0) {
#pragma unroll
                for (int q = 0; q < KB / 16; ++q)
                    *reinterpret_cast<float4 *>(dst + 4 * q) =
                        float4{pc[4 * q] * w, pc[4 * q + 1] * w, pc[4 * q + 2] * w, pc[4 * q + 3] * w};
            } else if constexpr (SPLIT) {
#pragma unroll
                for (int q = 0; q < KB / 32; ++q) {
                    uint32_t hi[4], lo[4];
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const float v0 = pc[8 * q + 2 * h] * w, v1 = pc[8 * q + 2 * h + 1] * w;
                        hi[h] = lpack_bf16(v0, v1);
                        lo[h] = lpack_bf16(v0 - __uint_as_float(hi[h] << 16), v1 - __uint_as_float(hi[h] & 0xffff0000u));
                    }
                    *reinterpret_cast<uint4 *>(dst + 8 * q) = uint4{hi[0], hi[1], hi[2], hi[3]};
                    *reinterpret_cast<uint4 *>(dst + Rp * XLD + 8 * q) = uint4{lo[0], lo[1], lo[2], lo[3]};
                }
            } else {
#pragma unroll
                for (int q = 0; q < KB / 32; ++q)
                    *reinterpret_cast<uint4 *>(dst + 8 * q) =
                        uint4{lpack_bf16(pc[8 * q] * w, pc[8 * q + 1] * w), lpack_bf16(pc[8 * q + 2] * w, pc[8 * q + 3] * w),
                              lpack_bf16(pc[8 * q + 4] * w, pc[8 * q + 5] * w), lpack_bf16(pc[8 * q + 6] * w, pc[8 * q + 7] * w)};
            }
        }
        __syncthreads();
        // refill b from the frontier (past the range it repeats the last index: loaded, never
        // consumed)
        load_tile(b, fc);
        __builtin_amdgcn_sched_barrier(0);  // the refill's loads stay ahead of the MFMAs
        if (!valid) return;
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
                for (int rb = 0; rb < 2; ++rb) {
                    if constexpr (F8T) {
                        // ds_read_b64_tr_b8: in each 16-lane group, lane l receives byte l & 7
                        // of the 8-byte pieces addressed by lanes 2j + (l >> 3), j = 0..7.  Lane
                        // 2j + h of group lg addresses fast index k = 32 q + 8 lg + j, rows
                        // 32 wave + 16 rb + 8 h .. + 7, so lane (lm, lg) gets X(row lm, k =
                        // 32 q + 8 lg + j) for j = 0..7: the A operand, widened here to BF16.
                        const int k = 32 * q + 8 * lg + (lm >> 1);
                        const int gr = 2 * wave + rb;  // 16-row granule
                        const unsigned char *src = reinterpret_cast<const unsigned char *>(xs) + k * 128 +
                                                   16 * (gr ^ ((k >> 1) & 7)) + 8 * (lm & 1);
                        typedef int lv2i __attribute__((ext_vector_type(2)));
                        const lv2i w = __builtin_amdgcn_ds_read_tr8_b64_v2i32(
                            (__attribute__((address_space(3))) lv2i *)(src));
                        av[rb] = __builtin_bit_cast(lbf16x8, lfp8x8_to_bf16(uint2{(uint32_t)w.x, (uint32_t)w.y}));
                    } else {
                        av[rb] = __builtin_bit_cast(
                            lbf16x8, *reinterpret_cast<const uint4 *>(xs + (32 * wave + 16 * rb + lm) * XLD + 32 * q + 8 * lg));
                    }
                }
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    bv[nb] = __builtin_bit_cast(
                        lbf16x8, *reinterpret_cast<const uint4 *>(bs + (16 * nb + lm) * XLD + 32 * q + 8 * lg));
#pragma unroll
                for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[rb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[rb], bv[nb], acc[rb][nb], 0, 0, 0);
                if constexpr (SPLIT) {
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        bv[nb] = __builtin_bit_cast(lbf16x8, *reinterpret_cast<const uint4 *>(
                                                                 bs + Rp * XLD + (16 * nb + lm) * XLD + 32 * q + 8 * lg));
#pragma unroll
                    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                        for (int nb = 0; nb < NB; ++nb)
                            acc[rb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[rb], bv[nb], acc[rb][nb], 0, 0, 0);
                }
            }
        }
    };

    const long nsl = sl1 - sl0;
    TileBuf A, B;
    for (long fc = fc0; fc < fc1 && nsl > 0; ++fc) {
#pragma unroll
        for (int g = 0; g < KB / 32; ++g) {
            // rows past the panel's padding (a multiple of 64) are zero
            const long i = fc * KB + kr_k0 + 8 * g;
            const bool pin = kr_on && i < prows;
            const float *pp = a.fpanel + (((pin ? i : 0) >> 3) * Rp + (kr_on ? kr_c : 0)) * 8;
            float4 u0 = *reinterpret_cast<const float4 *>(pp);
            float4 u1 = *reinterpret_cast<const float4 *>(pp + 4);
            if (!pin) u0 = u1 = float4{0.f, 0.f, 0.f, 0.f};
            pc[8 * g + 0] = u0.x; pc[8 * g + 1] = u0.y; pc[8 * g + 2] = u0.z; pc[8 * g + 3] = u0.w;
            pc[8 * g + 4] = u1.x; pc[8 * g + 5] = u1.y; pc[8 * g + 6] = u1.z; pc[8 * g + 7] = u1.w;
        }
        const bool full = row0 + 128 <= a.In && (fc + 1) * KB <= a.If;
        frontier_at(sl0, nsl);
        load_tile(A, fc);
        load_tile(B, fc);
        for (long s = 0; s < nsl; s += 2) {
            step(A, fc, full, true);
            step(B, fc, full, s + 1 < nsl);
        }
    }
    // this wave's 32 rows of the partial: lane holds rows 4 lg + v of each 16-row block
    // (TR other than FP8: LDS row j holds tile row RL (j % RG) + j / RG)
    float *dst = a.partial + (split * a.rows_pad + row0) * Rp;
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = 32 * wave + 16 * rb + 4 * lg + v;
            const int tr = (TR && !F8T) ? LTr<ST>::RL * (j % LTr<ST>::RG) + j / LTr<ST>::RG : j;
            float sc = 1.0f;
            if constexpr (TR && ST == 2) {
                const long r = row0 + tr;
                sc = (r < a.In) ? a.row_scale[r] : 0.0f;
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) dst[tr * Rp + 16 * nb + lm] = acc[rb][nb][v] * sc;
        }
}

// covered for rank <= 64 and orders 3 and 4 (at most two slow modes); mode 0 needs whole
// row groups (In a multiple of 4 rows for FP32, 8 for BF16, 16 for FP8)
bool mttkrp_lds_covers(int order, int mode, long rank, long In, int dtype) {
    if (rank > 64 || order < 3 || order > 4) return false;
    if (mode == 0) return In % ((dtype == TDX_F32) ? 4 : (dtype == TDX_FP8_E4M3) ? 16 : 8) == 0;
    return true;
}

int mttkrp_lds_launch(const MttkrpPlan &p, const void *X, const float *const *panels, const float *row_scale,
                      float *partial, hipStream_t s, const int *skip) {
    LArgs a{};
    a.X = X;
    for (int m = 0; m < 4; ++m) a.panel[m] = (m < p.order) ? panels[m] : nullptr;
    const int fm = (p.mode == 0) ? 1 : 0;
    a.fpanel = panels[fm];
    a.row_scale = row_scale;
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
    a.If = p.dims[fm];
    a.Rp = p.Rp;
    a.rowtiles = p.rowtiles;
    a.rows_pad = p.rows_pad;
    a.S = p.ksteps;             // slow indices in total
    a.sper = p.ksteps_per_split;
    a.cps = p.lds_cps;
    a.nss = p.lds_nss;
    a.ncg = p.nsplit / p.lds_nss;
    a.nslow = 0;
    for (int m = 1; m < p.order; ++m)
        if (m != p.mode && m != fm) a.slow[a.nslow++] = m;
    const int grid = p.rowtiles * p.nsplit;
    const int stt = (p.dtype == TDX_F32) ? 0 : (p.dtype == TDX_BF16) ? 1 : 2;
#define TDX_LDS_LAUNCH(ST_, TR_, SP_)                                                  \
    switch (p.Rp / 16) {                                                               \
        case 1: k_mttkrp_lds<ST_, 1, TR_, SP_><<<grid, 256, 0, s>>>(a); break;         \
        case 2: k_mttkrp_lds<ST_, 2, TR_, SP_><<<grid, 256, 0, s>>>(a); break;         \
        case 3: k_mttkrp_lds<ST_, 3, TR_, SP_><<<grid, 256, 0, s>>>(a); break;         \
        default: k_mttkrp_lds<ST_, 4, TR_, SP_><<<grid, 256, 0, s>>>(a); break;        \
    }
    if (p.split && stt != 1) return record(TDX_ERR_UNSUPPORTED);
    if (p.mode == 0) {
        if (stt == 0) { TDX_LDS_LAUNCH(0, true, false) }
        else if (stt == 1 && p.split) { TDX_LDS_LAUNCH(1, true, true) }
        else if (stt == 1) { TDX_LDS_LAUNCH(1, true, false) }
        else { TDX_LDS_LAUNCH(2, true, false) }
    } else {
        if (stt == 0) { TDX_LDS_LAUNCH(0, false, false) }
        else if (stt == 1 && p.split) { TDX_LDS_LAUNCH(1, false, true) }
        else if (stt == 1) { TDX_LDS_LAUNCH(1, false, false) }
        else { TDX_LDS_LAUNCH(2, false, false) }
    }
#undef TDX_LDS_LAUNCH
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

}  // namespace tdx
