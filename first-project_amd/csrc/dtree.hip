// dtree.hip -- the shared pass of an order-3 CP-ALS sweep (dimension tree), gfx950.
//
// Within a sweep (modes 0, 1, 2 in order) the mode-1 MTTKRP uses (A_new, C) and the mode-2
// MTTKRP uses (A_new, B_new).  Both contract X with A_new over i0 first:
//   T[i2][i1][r] = sum_{i0} X[i0, i1, i2] A_new[i0][r]
//   M1[i1][r]    = sum_{i2} T[i2][i1][r] C[i2][r]          (C is current during mode 1)
//   M2[i2][r]    = sum_{i1} T[i2][i1][r] B_new[i1][r]
// so one pass over X (k_ttm_m1: T and the mode-1 partials in the same pass) and one pass over
// T (k_m2, I1 I2 Rp BF16 values instead of I0 I1 I2 elements) replace two passes over X.  This is
// the dimension-tree MTTKRP of Phan, Tichavsky and Cichocki ("Fast alternating LS
// algorithms for high order CANDECOMP/PARAFAC tensor factorizations", IEEE TSP 61(19),
// 2013), restricted to order 3.  The arithmetic is that of the fused MTTKRP reassociated:
// the reference computes X(n) (KR product) per mode (src/cp_als.cu:46-52, 66-69, 85-88).
//
// k_ttm_m1 (BF16 or FP8 storage):
//   * a workgroup (8 or 4 waves, dt_waves) owns 32 i1 per wave and a range of i2; A_new (the mode-0
//     panel, pre-scaled by the FP8 slice scales) is converted once into BF16 MFMA operands
//     in LDS (I0 Rp 2 bytes, 128 KB at I0 = 1024, Rp = 64), in the exact per-lane order the
//     MFMA reads (one conflict-free 16-byte read per lane per operand);
//   * each wave computes T tiles (32 rows i1 x Rp) with v_mfma_f32_16x16x32_bf16: the A
//     operand is X read straight from HBM into registers, the B operand A_new from LDS.
//     The contraction order over i0 is free, so each lane owns 32 contiguous bytes of a
//     column per "run" (two k-steps of BF16, four of FP8; FP8 widened exactly to BF16 with
//     v_cvt_scalef32_pk_bf16_fp8).  Each 16-byte load instruction reads 8 whole 128-byte
//     lines (8 columns), and one DPP lane exchange (lane ^ 8) turns the pair of loads into
//     the operand order, where a load of 16 columns x 64 bytes each (half lines) streamed at
//     ~4.3 TB/s against ~6.2-6.7 for whole lines (513 -> 460 us at 1024^3 BF16 R 64);
//   * chunks of 4 (BF16) / 8 (FP8) k-steps (8 KB of X per wave) rotate through two (8 waves)
//     or four (4 waves) register buffers across the i2 loop, the others in flight while one is
//     multiplied
//     (nontemporal loads; every load is issued unconditionally, the chunks past the end
//     re-read a valid one), and after the last chunk of a column i2 the tile is multiplied
//     into the mode-1 partial by C[i2][:] (from LDS, FP32) and stored to T after the next
//     column's first chunk: T is BF16 (round to nearest even) in the MFMA output order, so a
//     lane's 16 bytes per row block go out as they are and each store instruction writes 1 KB
//     contiguous (NRB stores per tile; dt_t_index).  FP32 T written row-major with whole
//     lines per store took ~90 us of a 461 us pass at 1024^3 R64, half of it the write
//     traffic, half the issue of the stores;
//   * the mode-1 partial of the workgroup's i2 range goes to partial[split][i1][Rp], the
//     layout k_solve sums (in split order, FP64).
// k_m2: a wave per (kM2Slabs slabs i2, range of tiles); per tile a lane reads one 16-byte piece
//   of T per slab and row block (8 BF16: four consecutive i1 of two 16-column halves, one r)
//   and the matching rows of B_new once for the wave's slabs (one 16-byte panel load per
//   half), FP32 FMAs, the four lane groups added at the end; partial[split][i2][Rp].  With one
//   slab per wave two of every three load instructions were panel loads: 41.7 us at 1024^3
//   R64 against 26.7 with two slabs (four slabs: the same time, 130 VGPRs in k_m2_vinv).
#include "mttkrp.hpp"
#include "vinv.hpp"

namespace tdx {

typedef float dt_f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned dt_u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 dt_bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 dt_bf16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t dt_pack(float a, float b) {
    dt_bf16x2 v;
    v.x = (__bf16)a;  // round to nearest even (v_cvt_pk_bf16_f32)
    v.y = (__bf16)b;
    return __builtin_bit_cast(uint32_t, v);
}

// 4 FP8 E4M3 codes -> 4 BF16, exact (scale 1)
__device__ __forceinline__ uint2 dt_fp8x4(uint32_t q) {
    return uint2{__builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(q, 1.0f, false)),
                 __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(q, 1.0f, true))};
}

// Lanes l and l ^ 8 (DPP row_ror:8 within 16 lanes; hi: bit 3 of l set) exchange halves:
// p' = hi ? q(l ^ 8) : p,  q' = hi ? q : p(l ^ 8).  Written with update_dpp and an explicit
// old value: mov_dpp applied to the elements of a vector was merged into one exchange for all
// elements by this compiler.
__device__ __forceinline__ void swap_halves(uint32_t &p, uint32_t &q, bool hi) {
    const int s = (int)(hi ? p : q);
    const uint32_t x = (uint32_t)__builtin_amdgcn_update_dpp(s, s, 0x128, 0xf, 0xf, false);
    p = hi ? x : p;
    q = hi ? q : x;
}
__device__ __forceinline__ void swap_halves(uint4 &p, uint4 &q, bool hi) {
    swap_halves(p.x, q.x, hi);
    swap_halves(p.y, q.y, hi);
    swap_halves(p.z, q.z, hi);
    swap_halves(p.w, q.w, hi);
}

// ES: bytes per element; KPL: k-steps (32 i0 each) per 32-byte lane run; CH: k-steps per chunk
// (two runs: 8 KB of X per wave)
template <int ST> struct DtTile;
template <> struct DtTile<1> { static constexpr int ES = 2, KPL = 2, CH = 4; };
template <> struct DtTile<2> { static constexpr int ES = 1, KPL = 4, CH = 8; };

struct DtArgs {
    const void *X;
    const float *panelA;  // mode-0 panel (FP8: scaled by the slice scales)
    const float *C;       // mode-2 factor, column-major I2 x R
    uint16_t *T;          // BF16, tile layout (see dt_t_index)
    float *partial;       // nsplit1 x I1 x Rp
    long I0, I1, I2;
    int Rp, R;
    int i2per, blocks_i1;
};

// LDS: A operands ((I0/32) NRB 1 KB) and the C block (i2per Rp floats)
inline size_t dtree_lds_bytes(long I0, int nrb, int i2per, int Rp) {
    return (size_t)(I0 / 32) * nrb * 1024 + (size_t)i2per * Rp * 4;
}

// waves per workgroup of the pass over X by row blocks: 8 waves (two per SIMD, two chunk
// buffers, 190-212 VGPRs) sharing the A operands in LDS, or 4 waves with four buffers.
// Measured at 1024^3 (same box): 8 waves 3% faster at Rp 64, 1% at Rp 16, equal at Rp 48,
// 3% slower at Rp 32.
__host__ __device__ constexpr int dt_waves(int nrb) { return (nrb == 2 || nrb == 3) ? 4 : 8; }

template <int ST, int NRB, int NCH>
__global__ __launch_bounds__(64 * dt_waves(NRB), 1) void k_ttm_m1(DtArgs a) {
    constexpr int kDtWaves = dt_waves(NRB), kDtThreads = 64 * kDtWaves;
    constexpr int ES = DtTile<ST>::ES, KPL = DtTile<ST>::KPL, CH = DtTile<ST>::CH, RUNS = CH / KPL;
    constexpr int NB = kDtWaves == 8 ? 2 : (NCH < 4 ? NCH : 4);  // chunk buffers; NCH % NB == 0
    static_assert(NCH >= 2 && NCH % NB == 0, "NCH: chunks per column");
    extern __shared__ __align__(16) unsigned char dsm[];
    uint4 *aop = reinterpret_cast<uint4 *>(dsm);  // (I0/32) * NRB * 64 operands of 16 bytes
    const long nks = a.I0 / 32;
    float *cs = reinterpret_cast<float *>(dsm + nks * NRB * 1024);  // i2per x Rp
    const int tid = threadIdx.x, lane = tid & 63, lm = lane & 15, g = lane >> 4;
    const int wave = tid >> 6;
    const int bi = blockIdx.x % a.blocks_i1, split = blockIdx.x / a.blocks_i1;
    const long i2s = (long)split * a.i2per;
    const long i2e = (i2s + a.i2per < a.I2) ? i2s + a.i2per : a.I2;
    const int len = (int)(i2e - i2s);  // >= 1: the grid has ceil(I2 / i2per) splits
    // A operand (k-step ks, row block rb, lane l): rows r = 16 rb + l % 16, the 8 i0 of lane
    // group l / 16 within the run ks / KPL, step ks % KPL
    // (4 entries per thread in flight, the ones past nent repeat
    // the last entry's loads and are not stored)
    const long nent = nks * NRB * 64;
    for (long e0 = tid; e0 < nent; e0 += 4 * kDtThreads) {
        float4 p[4][2];
        long dst[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long e = (e0 + (long)kDtThreads * u < nent) ? e0 + (long)kDtThreads * u : nent - 1;
            const int l = (int)(e & 63);
            const long t = e >> 6;
            const int rb = (int)(t % NRB);
            const long ks = t / NRB;
            const int r = 16 * rb + (l & 15), gg = l >> 4;
            const long i0 = 32L * KPL * (ks / KPL) + 8L * KPL * gg + 8L * (ks % KPL);
            const float4 *pp = reinterpret_cast<const float4 *>(a.panelA + (((i0 >> 3) * a.Rp + r) << 3));
            p[u][0] = pp[0];
            p[u][1] = pp[1];
            dst[u] = e;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + (long)kDtThreads * u < nent)
                aop[dst[u]] = uint4{dt_pack(p[u][0].x, p[u][0].y), dt_pack(p[u][0].z, p[u][0].w),
                                    dt_pack(p[u][1].x, p[u][1].y), dt_pack(p[u][1].z, p[u][1].w)};
    }
    // C block: consecutive threads read consecutive i2 of one column (coalesced)
    for (int e = tid; e < len * a.Rp; e += kDtThreads) {
        const int r = e / len, il = e % len;
        cs[il * a.Rp + r] = r < a.R ? a.C[(i2s + il) + (long)r * a.I2] : 0.0f;
    }
    __syncthreads();  // the only barrier
    const long i1w = (long)bi * (32 * kDtWaves) + 32 * wave;  // first column of this wave's 32
    if (i1w >= a.I1) return;                      // I1 % 32 == 0: whole tiles only

    const unsigned char *X = static_cast<const unsigned char *>(a.X);
    const long colb = a.I0 * ES, slabb = a.I1 * colb;  // bytes per column i1, per slab i2
    constexpr long chb = (long)RUNS * 128;             // bytes of one chunk of a column
    const long total = (long)len * NCH;
    // Loads in line order: the two 16-byte loads of a lane's 32-byte run read columns
    // 16c + (lm & 7) (first) and 16c + 8 + (lm & 7) (second) at bytes 32g + 16 (lm >> 3) of the
    // run, so each load instruction reads 8 whole 128-byte lines; one lane-bit-3 exchange
    // (swap_halves) then gives each lane its operand-order 32 bytes of column 16c + lm.
    const bool hi8 = (lm & 8) != 0;
    const unsigned char *xl[2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
        xl[c] = X + (i2s * a.I1 + i1w + 16 * c + (lm & 7)) * colb + 8L * KPL * g * ES + 16 * (lm >> 3);
    const long last_off = (long)(len - 1) * slabb;  // re-read target past the end (valid)

    dt_f32x4 acc[NRB][2], m1[NRB][2];
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            acc[rb][c] = dt_f32x4{0.f, 0.f, 0.f, 0.f};
            m1[rb][c] = dt_f32x4{0.f, 0.f, 0.f, 0.f};
        }

    typedef uint4 Buf[2][RUNS][2];
    auto issue = [&](Buf &b, long off) {
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int j = 0; j < RUNS; ++j) {
                const unsigned char *p = xl[c] + off + j * 128;
                b[c][j][0] = ld_stream(reinterpret_cast<const uint4 *>(p));
                b[c][j][1] = ld_stream(reinterpret_cast<const uint4 *>(p + 8 * colb));
            }
    };
    auto compute = [&](Buf &b, int kc, int zoff) {
        const uint4 *ab = aop + zoff + kc * (CH * NRB * 64) + lane;
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int j = 0; j < RUNS; ++j) swap_halves(b[c][j][0], b[c][j][1], hi8);
#pragma unroll
        for (int j = 0; j < RUNS; ++j)
#pragma unroll
            for (int h = 0; h < KPL; ++h) {
                const int ksl = j * KPL + h;  // k-step within the chunk
                dt_bf16x8 av[NRB], bv[2];
#pragma unroll
                for (int rb = 0; rb < NRB; ++rb) av[rb] = __builtin_bit_cast(dt_bf16x8, ab[(ksl * NRB + rb) * 64]);
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    if constexpr (ST == 1) {
                        bv[c] = __builtin_bit_cast(dt_bf16x8, b[c][j][h]);
                    } else {
                        const uint4 &w = b[c][j][h >> 1];
                        const uint32_t q0 = (h & 1) ? w.z : w.x, q1 = (h & 1) ? w.w : w.y;
                        const uint2 lo = dt_fp8x4(q0), hi = dt_fp8x4(q1);
                        bv[c] = __builtin_bit_cast(dt_bf16x8, uint4{lo.x, lo.y, hi.x, hi.y});
                    }
                }
#pragma unroll
                for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                        acc[rb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[c], av[rb], acc[rb][c], 0, 0, 0);
            }
    };
    // finished tile (rows i1 = i1w + 16 c + 4 g + v, columns r = 16 rb + lm): into the mode-1
    // partial now, packed to BF16 (round to nearest even) for its T store
    uint4 tpk[NRB];
    auto take = [&](long i2l) {
        const float *cp = cs + i2l * a.Rp + lm;
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb) {
            const float cv = cp[16 * rb];
            tpk[rb] = uint4{dt_pack(acc[rb][0][0], acc[rb][0][1]), dt_pack(acc[rb][0][2], acc[rb][0][3]),
                            dt_pack(acc[rb][1][0], acc[rb][1][1]), dt_pack(acc[rb][1][2], acc[rb][1][3])};
#pragma unroll
            for (int c = 0; c < 2; ++c) {
#pragma unroll
                for (int v = 0; v < 4; ++v) m1[rb][c][v] = fmaf(acc[rb][c][v], cv, m1[rb][c][v]);
                acc[rb][c] = dt_f32x4{0.f, 0.f, 0.f, 0.f};
            }
        }
    };
    // T tile of this wave and column: NRB stores of 16 bytes per lane, each 1 KB contiguous
    auto store_t = [&](long i2l) {
        unsigned char *tp = reinterpret_cast<unsigned char *>(a.T) +
                            ((i2s + i2l) * (a.I1 / 32) + i1w / 32) * 64L * a.Rp + 16 * lane;
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
            __builtin_nontemporal_store(dt_u32x4{tpk[rb].x, tpk[rb].y, tpk[rb].z, tpk[rb].w},
                                        reinterpret_cast<dt_u32x4 *>(tp + 1024 * rb));
    };

    // The chunk sequence runs column after column (NCH chunks each); chunk q lives in buffer
    // q % NB, and the loads of chunks q+1 .. q+NB-1 are in flight while chunk q is multiplied.
    // Columns are unrolled over their chunks (the loop body has no branches) and a column's
    // T store is issued after the next column's first chunk, so on every path into the loop
    // the stores are older than the loads in flight and the compiler's load-counter waits
    // keep NB-1 chunks outstanding (stores issued after the prefetch made them drain it at
    // every column).  Each wave walks a column's chunks from its own rotation (the order of
    // the sum over i0 is free).  Chunks past the end re-read the last column's first chunk
    // (a valid address, never multiplied).
    const int rot = (int)((blockIdx.x * (unsigned)kDtWaves + (unsigned)wave) & (NCH - 1));
    long qi = 0, i2_i = 0;
    int kc_i = 0;
    auto next_off = [&]() {
        const long o = qi < total ? i2_i * slabb + (long)((kc_i + rot) & (NCH - 1)) * chb : last_off;
        ++qi;
        const bool w = kc_i == NCH - 1;
        i2_i += w ? 1 : 0;
        kc_i = w ? 0 : kc_i + 1;
        return o;
    };
    Buf b[NB];
#pragma unroll
    for (int u = 0; u < NB - 1; ++u) issue(b[u], next_off());
#pragma unroll
    for (int kc = 0; kc < NCH; ++kc) {  // column 0
        issue(b[(kc + NB - 1) % NB], next_off());
        __builtin_amdgcn_sched_barrier(0);
        compute(b[kc % NB], (kc + rot) & (NCH - 1), 0);
        __builtin_amdgcn_sched_barrier(0);
    }
    take(0);
    for (int i2l = 1; i2l < len; ++i2l) {
        // the A operands of a chunk are the same for every column: an opaque zero offset
        // keeps their LDS reads inside the loop (hoisted, they held every chunk's operands
        // in registers at once)
        int zoff = 0;
        asm volatile("" : "+s"(zoff));
#pragma unroll
        for (int kc = 0; kc < NCH; ++kc) {
            issue(b[(kc + NB - 1) % NB], next_off());
            __builtin_amdgcn_sched_barrier(0);  // the loads stay ahead of this chunk's MFMAs
            compute(b[kc % NB], (kc + rot) & (NCH - 1), zoff);
            __builtin_amdgcn_sched_barrier(0);
            if (kc == 0) {
                store_t(i2l - 1);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        take(i2l);
    }
    store_t(len - 1);
    // mode-1 partial of this split
#pragma unroll
    for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                a.partial[((long)split * a.I1 + i1w + 16 * c + 4 * g + v) * a.Rp + 16 * rb + lm] = m1[rb][c][v];
}

// slabs of T per wave in the pass over T: each panel load serves that many slabs
static constexpr int kM2Slabs = 2;

// M2 partials: wave w of block b takes the kM2Slabs slabs i2 = kM2Slabs (4 b + w) + j and the T
// tiles [split tper, ...) of each; lane (lm, g) holds, per slab, tile and row block rb, the BF16
// values of i1 = 32 tile + 16 c + 4 g + v (c = 0, 1; v = 0..3) and r = 16 rb + lm (one 16-byte
// load, the layout the pass over X stores) and multiplies them by B_new[i1][r] (one 16-byte
// panel load per c, four consecutive rows of one column, shared by the wave's slabs); the four
// lane groups g are added by lane exchanges at the end.  Slabs past I2 re-read slab I2 - 1 and
// store nothing.  blk is the block's index among the pass's blocks.
template <int NRB>
__device__ __forceinline__ void m2_block(const uint16_t *__restrict__ T, const float *__restrict__ panelB,
                                         float *__restrict__ partial, long I1, long I2, int Rp, int tper, int nb2,
                                         int blk) {
    constexpr int J = kM2Slabs;
    const int tid = threadIdx.x, lane = tid & 63, lm = lane & 15, g = lane >> 4, wave = tid >> 6;
    const int split = blk / nb2, b = blk % nb2;
    const long i2b = (long)J * (4L * b + wave);
    if (i2b >= I2) return;  // no barriers in this function
    const long ntile = I1 / 32;
    const long t0 = (long)split * tper, t1 = (t0 + tper < ntile) ? t0 + tper : ntile;
    const unsigned char *tb[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const long i2 = (i2b + j < I2) ? i2b + j : I2 - 1;
        tb[j] = reinterpret_cast<const unsigned char *>(T) + i2 * ntile * 64L * Rp + 16 * lane;
    }
    float acc[J][NRB];
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb) acc[j][rb] = 0.0f;
    for (long tw = t0; tw < t1; ++tw) {
        dt_u32x4 w[J][NRB];
        float4 bb[NRB][2];
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int rb = 0; rb < NRB; ++rb)
                w[j][rb] = __builtin_nontemporal_load(
                    reinterpret_cast<const dt_u32x4 *>(tb[j] + tw * 64L * Rp + 1024 * rb));
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const long i1b = 32 * tw + 16 * c + 4 * g;
                bb[rb][c] = *reinterpret_cast<const float4 *>(panelB + (((i1b >> 3) * Rp + 16 * rb + lm) << 3) +
                                                              4 * (g & 1));
            }
#pragma unroll
        for (int j = 0; j < J; ++j)
#pragma unroll
            for (int rb = 0; rb < NRB; ++rb) {
                const unsigned q[4] = {w[j][rb].x, w[j][rb].y, w[j][rb].z, w[j][rb].w};
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    const float4 &y = bb[rb][c];
                    const unsigned p0 = q[2 * c], p1 = q[2 * c + 1];
                    float &s = acc[j][rb];
                    s = fmaf(__uint_as_float(p0 << 16), y.x, s);
                    s = fmaf(__uint_as_float(p0 & 0xffff0000u), y.y, s);
                    s = fmaf(__uint_as_float(p1 << 16), y.z, s);
                    s = fmaf(__uint_as_float(p1 & 0xffff0000u), y.w, s);
                }
            }
    }
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
        for (int rb = 0; rb < NRB; ++rb) {
            acc[j][rb] += __shfl_xor(acc[j][rb], 16);
            acc[j][rb] += __shfl_xor(acc[j][rb], 32);
        }
    if (g == 0) {
#pragma unroll
        for (int j = 0; j < J; ++j)
            if (i2b + j < I2) {
#pragma unroll
                for (int rb = 0; rb < NRB; ++rb)
                    partial[((long)split * I2 + i2b + j) * Rp + 16 * rb + lm] = acc[j][rb];
            }
    }
}

template <int NRB>
__global__ __launch_bounds__(256) void k_m2(const uint16_t *__restrict__ T, const float *__restrict__ panelB,
                                            float *__restrict__ partial, long I1, long I2, int Rp, int tper,
                                            int nb2) {
    m2_block<NRB>(T, panelB, partial, I1, I2, Rp, tper, nb2, blockIdx.x);
}

// The pass over T with the mode-2 inverse (G0 * G1)^{-1} in the same launch: block 0
// computes it (as k_vinv), blocks 1.. are k_m2's.  The inverse block comes first in dispatch
// order, so it runs beside the pass instead of after it, without a fork and a join; the solve
// after this kernel needs nothing else.  G1 is summed by the launch before this one: summed
// here by leading blocks, with an atomic ticket choosing the block that inverts, the sum and
// the ticket took ~17 us at 1024^3 R64 (memory round trips behind the pass's loads) before
// the inverse could start.
template <int NRB, int NS>
__global__ __launch_bounds__(256) void k_m2_vinv(const uint16_t *__restrict__ T, const float *__restrict__ panelB,
                                                 float *__restrict__ partial, long I1, long I2, int Rp, int tper,
                                                 int nb2, const double *G0, const double *G1, int R,
                                                 double *__restrict__ Vinv, int *__restrict__ status,
                                                 double *__restrict__ fb_ws) {
    if (blockIdx.x > 0) {
        m2_block<NRB>(T, panelB, partial, I1, I2, Rp, tper, nb2, (int)blockIdx.x - 1);
        return;
    }
    const double *G[4] = {G0, G1, G0, G0};
    vinv_block<NS>(G, 3, 2, R, Vinv, status, fb_ws);
}

// ---------------------------------------------------------------------------
// host side

// a == nullptr: set the kernel's dynamic LDS limit; otherwise launch it
template <int ST, int NRB, int NCH>
static int m1_kernel(const DtreePlan &d, const DtArgs *a, hipStream_t s) {
    if (!a) {
        TDX_HIP(hipFuncSetAttribute((const void *)k_ttm_m1<ST, NRB, NCH>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)d.lds1));
        return TDX_OK;
    }
    k_ttm_m1<ST, NRB, NCH><<<d.blocks_i1 * d.nsplit1, 64 * dt_waves(NRB), d.lds1, s>>>(*a);
    return TDX_OK;
}

template <int ST, int NRB>
static int m1_by_nch(const DtreePlan &d, const DtArgs *a, hipStream_t s) {
    switch (d.nch) {
        case 2: return m1_kernel<ST, NRB, 2>(d, a, s);
        case 4: return m1_kernel<ST, NRB, 4>(d, a, s);
        case 8: return m1_kernel<ST, NRB, 8>(d, a, s);
        case 16: return m1_kernel<ST, NRB, 16>(d, a, s);
        default: return record(TDX_ERR_UNSUPPORTED);
    }
}

template <int ST>
static int m1_by_nrb(const DtreePlan &d, const DtArgs *a, hipStream_t s) {
    switch (d.nrb) {
        case 1: return m1_by_nch<ST, 1>(d, a, s);
        case 2: return m1_by_nch<ST, 2>(d, a, s);
        case 3: return m1_by_nch<ST, 3>(d, a, s);
        case 4: return m1_by_nch<ST, 4>(d, a, s);
        default: return record(TDX_ERR_UNSUPPORTED);
    }
}

static int m1_call(const DtreePlan &d, const DtArgs *a, hipStream_t s) {
    return d.st == 1 ? m1_by_nrb<1>(d, a, s) : m1_by_nrb<2>(d, a, s);
}

int dtree_plan(DtreePlan &d, int order, const long *dims, int dtype, long R, bool enabled) {
    d = DtreePlan();
    if (!enabled || order != 3 || R < 1 || R > 64) return TDX_OK;
    if (dtype != TDX_BF16 && dtype != TDX_FP8_E4M3) return TDX_OK;
    const int st = (dtype == TDX_BF16) ? 1 : 2;
    const long CH = (st == 1) ? DtTile<1>::CH : DtTile<2>::CH;
    d.st = st;
    d.R = (int)R;
    d.Rp = (int)(ceil_div(R, 16) * 16);
    d.nrb = d.Rp / 16;
    d.I0 = dims[0];
    d.I1 = dims[1];
    d.I2 = dims[2];
    if (d.I0 % (32 * CH) != 0 || d.I1 % 32 != 0 || d.I2 < 1) return TDX_OK;
    d.nch = (int)(d.I0 / (32 * CH));  // chunks per column: 2, 4, 8 or 16
    if (d.nch != 2 && d.nch != 4 && d.nch != 8 && d.nch != 16) return TDX_OK;
    const long cus = device_cus();
    d.blocks_i1 = (int)ceil_div(d.I1, 32L * dt_waves(d.nrb));
    long ns = cus / d.blocks_i1;
    if (ns < 1) ns = 1;
    if (ns > d.I2) ns = d.I2;
    d.i2per = (int)ceil_div(d.I2, ns);
    d.nsplit1 = (int)ceil_div(d.I2, (long)d.i2per);
    d.lds1 = dtree_lds_bytes(d.I0, d.nrb, d.i2per, d.Rp);
    if (d.lds1 > 160 * 1024) return TDX_OK;  // LDS per workgroup on gfx950
    // k_m2: one wave per (kM2Slabs slabs, tile range), about 16 waves per CU: nsplit2 ranges
    // of tper tiles
    const long ntile = d.I1 / 32;
    long n2 = ceil_div(16 * cus * kM2Slabs, d.I2);
    if (n2 > ntile) n2 = ntile;
    if (n2 < 1) n2 = 1;
    d.tper = (int)ceil_div(ntile, n2);
    d.nsplit2 = (int)ceil_div(ntile, (long)d.tper);
    d.nb2 = (int)ceil_div(d.I2, 4L * kM2Slabs);
    const int rc = m1_call(d, nullptr, nullptr);
    if (rc) return rc;
    d.on = true;
    return TDX_OK;
}

int dtree_m1_launch(const DtreePlan &d, const void *X, const float *panelA, const float *C, uint16_t *T,
                    float *partial, hipStream_t s) {
    if (!d.on) return record(TDX_ERR_UNSUPPORTED);
    if ((reinterpret_cast<uintptr_t>(X) & 15) != 0) return record(TDX_ERR_ARG);
    DtArgs a;
    a.X = X;
    a.panelA = panelA;
    a.C = C;
    a.T = T;
    a.partial = partial;
    a.I0 = d.I0;
    a.I1 = d.I1;
    a.I2 = d.I2;
    a.Rp = d.Rp;
    a.R = d.R;
    a.i2per = d.i2per;
    a.blocks_i1 = d.blocks_i1;
    const int rc = m1_call(d, &a, s);
    if (rc) return rc;
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

int dtree_m2_vinv_launch(const DtreePlan &d, const uint16_t *T, const float *panelB, float *partial,
                         const double *G0, const double *G1, int R, double *Vinv, int *status, double *fb_ws,
                         hipStream_t s) {
    if (!d.on || R <= 16 || R > kMaxR) return record(TDX_ERR_UNSUPPORTED);
    const int grid = 1 + d.nb2 * d.nsplit2;
#define TDX_M2V(NRB_, NS_)                                                                                     \
    k_m2_vinv<NRB_, NS_><<<grid, 256, 0, s>>>(T, panelB, partial, d.I1, d.I2, d.Rp, d.tper, d.nb2, G0, G1, R, \
                                              Vinv, status, fb_ws)
    // 16 < R <= 32: Rp 32 (NRB 2); 32 < R <= 64: Rp 48 or 64
    if (d.nrb == 2) TDX_M2V(2, 8);
    else if (d.nrb == 3) TDX_M2V(3, 16);
    else if (d.nrb == 4) TDX_M2V(4, 16);
    else return record(TDX_ERR_UNSUPPORTED);
#undef TDX_M2V
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

int dtree_m2_launch(const DtreePlan &d, const uint16_t *T, const float *panelB, float *partial, hipStream_t s) {
    if (!d.on) return record(TDX_ERR_UNSUPPORTED);
    const int grid = d.nb2 * d.nsplit2;
    switch (d.nrb) {
        case 1: k_m2<1><<<grid, 256, 0, s>>>(T, panelB, partial, d.I1, d.I2, d.Rp, d.tper, d.nb2); break;
        case 2: k_m2<2><<<grid, 256, 0, s>>>(T, panelB, partial, d.I1, d.I2, d.Rp, d.tper, d.nb2); break;
        case 3: k_m2<3><<<grid, 256, 0, s>>>(T, panelB, partial, d.I1, d.I2, d.Rp, d.tper, d.nb2); break;
        case 4: k_m2<4><<<grid, 256, 0, s>>>(T, panelB, partial, d.I1, d.I2, d.Rp, d.tper, d.nb2); break;
        default: return record(TDX_ERR_UNSUPPORTED);
    }
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

}  // namespace tdx

using namespace tdx;

extern "C" {

int tdx_mttkrp_dtree(const void *X, int dtype, const float *scales, const long *dims, const float *const *factors,
                     long rank, float *M1, float *M2) {
    if (!X || !dims || !factors || !factors[0] || !factors[1] || !factors[2] || !M1 || !M2)
        return record(TDX_ERR_ARG);
    if (dtype == TDX_FP8_E4M3 && !scales) return record(TDX_ERR_ARG);
    DtreePlan d;
    int st = dtree_plan(d, 3, dims, dtype, rank, true);
    if (st) return st;
    if (!d.on) return record(TDX_ERR_UNSUPPORTED);
    hipStream_t s = stream();
    float *panel[2] = {nullptr, nullptr};
    for (int m = 0; m < 2; ++m) {
        void *ws = nullptr;
        st = workspace(WS_PANEL0 + m, sizeof(float) * panel_floats(dims[m], d.Rp), &ws);
        if (st) return st;
        st = pack_panel(factors[m], dims[m], rank, d.Rp, (dtype == TDX_FP8_E4M3 && m == 0) ? scales : nullptr,
                        (float *)ws, s);
        if (st) return record(st);
        panel[m] = (float *)ws;
    }
    void *T = nullptr, *part = nullptr;
    st = workspace(WS_MISC0, d.t_bytes(), &T);
    if (st) return st;
    st = workspace(WS_PARTIAL, sizeof(float) * d.partial_floats(), &part);
    if (st) return st;
    st = dtree_m1_launch(d, X, panel[0], factors[2], (uint16_t *)T, (float *)part, s);
    if (st) return st;
    MttkrpPlan p1;
    p1.order = 3;
    for (int m = 0; m < 3; ++m) p1.dims[m] = dims[m];
    p1.rank = rank;
    p1.Rp = d.Rp;
    p1.mode = 1;
    p1.rows_pad = d.I1;
    p1.nsplit = d.nsplit1;
    st = mttkrp_reduce(p1, (const float *)part, M1, s);
    if (st) return record(st);
    st = dtree_m2_launch(d, (const uint16_t *)T, panel[1], (float *)part, s);
    if (st) return st;
    MttkrpPlan p2 = p1;
    p2.mode = 2;
    p2.rows_pad = d.I2;
    p2.nsplit = d.nsplit2;
    st = mttkrp_reduce(p2, (const float *)part, M2, s);
    return record(st);
}

// Times `iters` launches of each kernel of the shared pass (panels and workspace prepared
// once, one warm-up launch each) with hipEvents on the library stream; *ms_m1 / *ms_m2 receive
// the mean launch duration of k_ttm_m1 (the pass over X) and of k_m2 (the pass over T).
int tdx_bench_mttkrp_dtree(const void *X, int dtype, const float *scales, const long *dims,
                           const float *const *factors, long rank, int iters, float *ms_m1, float *ms_m2) {
    if (!X || !dims || !factors || !factors[0] || !factors[1] || !factors[2] || !ms_m1 || !ms_m2 || iters <= 0)
        return record(TDX_ERR_ARG);
    if (dtype == TDX_FP8_E4M3 && !scales) return record(TDX_ERR_ARG);
    DtreePlan d;
    int st = dtree_plan(d, 3, dims, dtype, rank, true);
    if (st) return st;
    if (!d.on) return record(TDX_ERR_UNSUPPORTED);
    hipStream_t s = stream();
    float *panel[2] = {nullptr, nullptr};
    for (int m = 0; m < 2; ++m) {
        void *ws = nullptr;
        st = workspace(WS_PANEL0 + m, sizeof(float) * panel_floats(dims[m], d.Rp), &ws);
        if (st) return st;
        st = pack_panel(factors[m], dims[m], rank, d.Rp, (dtype == TDX_FP8_E4M3 && m == 0) ? scales : nullptr,
                        (float *)ws, s);
        if (st) return record(st);
        panel[m] = (float *)ws;
    }
    void *T = nullptr, *part = nullptr;
    st = workspace(WS_MISC0, d.t_bytes(), &T);
    if (st) return st;
    st = workspace(WS_PARTIAL, sizeof(float) * d.partial_floats(), &part);
    if (st) return st;
    auto chk = [](hipError_t e, const char *what) {
        return e == hipSuccess ? TDX_OK : record_hip(e, what, __FILE__, __LINE__);
    };
    hipEvent_t e[3];
    for (int i = 0; i < 3; ++i) TDX_HIP(hipEventCreate(&e[i]));
    st = dtree_m1_launch(d, X, panel[0], factors[2], (uint16_t *)T, (float *)part, s);
    if (!st) st = dtree_m2_launch(d, (const uint16_t *)T, panel[1], (float *)part, s);
    if (!st) st = chk(hipEventRecord(e[0], s), "hipEventRecord");
    for (int i = 0; i < iters && !st; ++i)
        st = dtree_m1_launch(d, X, panel[0], factors[2], (uint16_t *)T, (float *)part, s);
    if (!st) st = chk(hipEventRecord(e[1], s), "hipEventRecord");
    for (int i = 0; i < iters && !st; ++i) st = dtree_m2_launch(d, (const uint16_t *)T, panel[1], (float *)part, s);
    if (!st) st = chk(hipEventRecord(e[2], s), "hipEventRecord");
    if (!st) st = chk(hipEventSynchronize(e[2]), "hipEventSynchronize");
    float t1 = 0.f, t2 = 0.f;
    if (!st) {
        (void)hipEventElapsedTime(&t1, e[0], e[1]);
        (void)hipEventElapsedTime(&t2, e[1], e[2]);
        *ms_m1 = t1 / iters;
        *ms_m2 = t2 / iters;
    }
    for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]);
    return record(st);
}

}  // extern "C"
