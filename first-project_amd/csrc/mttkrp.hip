// mttkrp.hip -- fused MTTKRP for gfx950 (CDNA4).
//
// M[i_n, r] = sum over all other indices of X[...] * prod_{m != n} U_m[i_m, r]
//
// The tensor streams from HBM exactly once per call; the Khatri-Rao rows are formed in
// registers from factor "panels" and never materialised (the reference writes a
// (prod I_m) x R KR matrix with a batched GEMM per mode, src/cp_als.cu:46,66,85, and for
// mode 2 also an unfolded copy of X, src/cp_als.cu:26-28).
//
// Contraction view.  The output mode n has extent In.  The fastest contracted mode f
// (mode 0, or mode 1 when n == 0) is walked in K steps of 16 (FP32) or 32 (BF16/FP8)
// indices; every other contracted ("slow") mode is fixed within a K step, so its factor
// rows fold into one per-column scalar w[r].  A K step therefore multiplies a
// (rows x KS) block of X by the (KS x R) block  panel_f[f0:f0+KS, :] * w.
//
// Two load shapes:
//  * case B (n > 0): X is contiguous along the contraction (mode 0).  Each lane loads 16 B
//    (8 BF16 / 4 FP32) or 8 B (8 FP8) straight into its MFMA A-operand slot.
//  * case A (n == 0): X is contiguous along the OUTPUT rows.  Each lane loads 8 rows for
//    8 consecutive K indices and transposes the 8x8 block inside its registers, so the
//    MFMA operand is built without LDS.
// MFMA: v_mfma_f32_16x16x32_bf16 (BF16 and FP8 storage, FP8 widened exactly to BF16) and
// v_mfma_f32_16x16x4_f32 (FP32), accumulation in FP32.  Each wave owns 128 (64 for FP32)
// output rows x Rp columns; the 4 waves of a workgroup walk interleaved K steps of one
// split, are summed in LDS in a fixed order, and the workgroup writes one FP32 partial.
// Partials are summed over splits in FP64 by the consumer (mttkrp_reduce or the ALS solve).
#include "mttkrp.hpp"
#include <stdlib.h>
#include <mutex>

namespace tdx {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

struct KArgs {
    const void *X;
    const float *panel[4];
    const float *row_scale;
    float *partial;
    const int *skip;  // when non-null and *skip != 0 the launch does nothing (partials reused)
    long stride[4];
    long dims[4];
    long In, If, FB, ksteps, kper, rows_pad;
    int order, mode, fmode, nslow, Rp, rowtiles;
    int slow[3];
};

__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
    bf16x2 v;
    v.x = (__bf16)a;
    v.y = (__bf16)b;
    return __builtin_bit_cast(uint32_t, v);
}

// 8 FP8 E4M3 codes (little-endian in two dwords) -> 8 BF16 packed in 4 dwords; exact
// (every E4M3 value is a BF16 value; v_cvt_scalef32_pk_bf16_fp8 with scale 1 converts two
// codes per instruction, bit-identical to widening through FP32 for all 256 codes)
__device__ __forceinline__ uint4 fp8x8_to_bf16x8(uint2 q) {
    uint4 o;
    o.x = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(q.x, 1.0f, false));
    o.y = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(q.x, 1.0f, true));
    o.z = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(q.y, 1.0f, false));
    o.w = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_scalef32_pk_bf16_fp8(q.y, 1.0f, true));
    return o;
}

__device__ __forceinline__ uint32_t u4_get(const uint4 &v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
}

// 16-bit element t of an 8-element vector held in a uint4
__device__ __forceinline__ uint32_t half_of(const uint4 &v, int t) {
    uint32_t w = u4_get(v, t >> 1);
    return (t & 1) ? (w >> 16) : (w & 0xffffu);
}

// element type and per-lane load of one K-step operand slice
template <int ST> struct XElem;
template <> struct XElem<0> { typedef float T; typedef float4 V; };    // 4 FP32 per 16 B
template <> struct XElem<1> { typedef uint16_t T; typedef uint4 V; };  // 8 BF16 per 16 B
template <> struct XElem<2> { typedef uint8_t T; typedef uint2 V; };   // 8 FP8 per 8 B

// keep the loaded bits where mask is all-ones, zero them where it is 0
__device__ __forceinline__ float4 vmask(float4 v, uint32_t m) {
    return float4{__uint_as_float(__float_as_uint(v.x) & m), __uint_as_float(__float_as_uint(v.y) & m),
                  __uint_as_float(__float_as_uint(v.z) & m), __uint_as_float(__float_as_uint(v.w) & m)};
}
__device__ __forceinline__ uint4 vmask(uint4 v, uint32_t m) { return uint4{v.x & m, v.y & m, v.z & m, v.w & m}; }
__device__ __forceinline__ uint2 vmask(uint2 v, uint32_t m) { return uint2{v.x & m, v.y & m}; }

// Loads of one K step.  Addresses are clamped into the tensor so every load is issued
// unconditionally; which loads were in range is kept as a bit mask and out-of-range values
// are zeroed only when the step is consumed (zeroing at issue time would make the wave wait
// for the loads right after issuing them).
template <int ST, bool CA, int T>
struct StepLoads {
    typedef typename XElem<ST>::T ET;
    typedef typename XElem<ST>::V VT;
    static constexpr int NL = (ST == 0) ? 4 : 8;  // loads per lane per K step
    VT v[NL];
    uint32_t ok;  // bit q: load q was in range
    __device__ __forceinline__ void issue(const ET *__restrict__ X, long slow_off, long fj, long If, long s1,
                                          const long *roff, const bool *rvalid) {
        ok = 0;
        if constexpr (CA) {
            // lane rows fixed (roff[0]); NL consecutive fast indices fj .. fj+NL-1
#pragma unroll
            for (int q = 0; q < NL; ++q) {
                const long j = fj + q;
                const bool in = rvalid[0] && (j < If);
                v[q] = ld_stream(reinterpret_cast<const VT *>(X + slow_off + (in ? j : 0) * s1 + roff[0]));
                ok |= (uint32_t)in << q;
            }
        } else {
            const bool fok = fj < If;
#pragma unroll
            for (int t = 0; t < NL; ++t) {
                v[t] = ld_stream(reinterpret_cast<const VT *>(X + slow_off + (fok ? fj : 0) + roff[t]));
                ok |= (uint32_t)(rvalid[t] && fok) << t;
            }
        }
    }
    // load q with out-of-range lanes zeroed
    __device__ __forceinline__ VT get(int q) const { return vmask(v[q], ((ok >> q) & 1u) ? 0xffffffffu : 0u); }
};

template <int ST, bool CA, int NB>
__global__ __launch_bounds__(256, (ST != 0 && NB >= 3) ? 1 : 2) void k_mttkrp(KArgs a) {
    constexpr int KS = (ST == 0) ? 16 : 32;
    constexpr int T = (ST == 0) ? 4 : 8;
    constexpr int RW = 16 * T;
    typedef typename XElem<ST>::T ET;
    __shared__ float red[RW * 64];
    if (a.skip && *a.skip) return;

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const int lm = lane & 15, lg = lane >> 4;
    const int rowtile = blockIdx.x % a.rowtiles;
    const long split = blockIdx.x / a.rowtiles;
    const long row0 = (long)rowtile * RW;
    const long kbeg = split * a.kper;
    const long kend = min(kbeg + a.kper, a.ksteps);
    const int Rp = a.Rp;
    const ET *__restrict__ X = reinterpret_cast<const ET *>(a.X);
    const long In = a.In, If = a.If, FB = a.FB;
    const long s1 = a.stride[1];
    const long smode = a.stride[a.mode];
    const float *__restrict__ pf = a.panel[a.fmode];

    // per-lane row offsets (elements), fixed for the whole K range
    long roff[T];
    bool rvalid[T];
    if constexpr (CA) {
        const long r = row0 + ((ST == 0) ? 4 : 8) * lm;
        rvalid[0] = r < In;
        roff[0] = rvalid[0] ? r : 0;
#pragma unroll
        for (int t = 1; t < T; ++t) { roff[t] = 0; rvalid[t] = false; }
    } else {
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const long o = row0 + 16 * t + lm;
            rvalid[t] = o < In;
            roff[t] = (rvalid[t] ? o : 0) * smode;
        }
    }
    const long lane_f = (ST == 0) ? 4 * lg : 8 * lg;  // this lane's K-slot offset inside a step

    f32x4 acc[T][NB];
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[t][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Slow-mode weights w[r] = prod of the slow factor rows for every slow index this
    // workgroup touches, precomputed into LDS (the reduction buffer, free until the end), so
    // the K loop issues only tensor and panel loads.
    const long sl_first = kbeg / FB;
    const long sl_count = (kend > kbeg) ? (kend - 1) / FB - sl_first + 1 : 0;
    float *wtab = red;  // sl_count * Rp floats, guaranteed to fit by the planner
    for (long e = threadIdx.x; e < sl_count * Rp; e += blockDim.x) {
        long rem = sl_first + e / Rp;
        const int r = (int)(e % Rp);
        float w = 1.0f;
        for (int q = 0; q < a.nslow; ++q) {
            const int m = a.slow[q];
            const long idx = rem % a.dims[m];
            rem /= a.dims[m];
            w *= a.panel[m][((idx >> 3) * Rp + r) * 8 + (idx & 7)];
        }
        wtab[e] = w;
    }
    __syncthreads();
    auto slow_state = [&](long sl, float *w, long &off) {
        long rem = sl;
        off = 0;
        for (int q = 0; q < a.nslow; ++q) {
            const int m = a.slow[q];
            const long idx = rem % a.dims[m];
            rem /= a.dims[m];
            off += idx * a.stride[m];
        }
        const float *wt = wtab + (sl - sl_first) * Rp + lm;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) w[nb] = wt[16 * nb];
    };

    // Software pipeline: PD register slots, each holding one K step's tensor loads, its
    // fast-block index and its slow-mode weights.  Loads for step i + PD - 1 are issued
    // before step i is multiplied, so PD - 1 steps are always in flight per wave.
    constexpr int PD = (ST != 0 && NB >= 3) ? 3 : 2;
    // the slot's factor-panel values are loaded together with its tensor loads, so that
    // waiting for one slot never waits for loads issued for later slots (vmcnt is in order)
    constexpr int PV = (ST == 0) ? 1 : 2;  // float4 panel values per column block
    struct Slot {
        StepLoads<ST, CA, T> v;
        float4 pan[NB][PV];
        float w[NB];
    };
    // each wave takes a contiguous quarter of the workgroup's K steps, so the two halves of
    // a cache line (consecutive K steps of a row in case B) are read by the same wave
    // back to back
    const long quarter = (kend - kbeg + 3) / 4;
    const long ks0 = kbeg + wave * quarter;
    const long ks1 = min(ks0 + quarter, kend);
    const long nsteps = (ks0 < ks1) ? ks1 - ks0 : 0;
    long f_fb = ks0 % FB, f_sl = ks0 / FB, f_off = 0;
    float f_w[NB];
    if (nsteps > 0) slow_state(f_sl, f_w, f_off);
    long issued = 0;
    auto issue = [&](Slot &sl) {
        sl.v.issue(X, f_off, f_fb * KS + lane_f, If, s1, roff, rvalid);
        {
            const long fj = f_fb * KS + lane_f;
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const float *pp = pf + ((fj >> 3) * Rp + 16 * nb + lm) * 8 + (fj & 7);
#pragma unroll
                for (int q = 0; q < PV; ++q) sl.pan[nb][q] = *reinterpret_cast<const float4 *>(pp + 4 * q);
            }
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) sl.w[nb] = f_w[nb];
        // advance the frontier; past the last step it stays put, so the surplus issues of
        // the pipeline tail re-read the last step's lines (cache hits, never consumed)
        if (++issued < nsteps) {
            f_fb += 1;
            bool moved = false;
            if (f_fb >= FB) { f_fb = 0; ++f_sl; moved = true; }
            if (moved) slow_state(f_sl, f_w, f_off);
        }
    };
    auto compute = [&](const Slot &sl) {
        if constexpr (ST == 0) {
            float kr[NB][4];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const float4 p = sl.pan[nb][0];
                kr[nb][0] = p.x * sl.w[nb];
                kr[nb][1] = p.y * sl.w[nb];
                kr[nb][2] = p.z * sl.w[nb];
                kr[nb][3] = p.w * sl.w[nb];
            }
            // A operand: row block t, K substep s (K slot lg <-> f0 + 4 lg + s)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    float xv;
                    if constexpr (CA) {
                        const float4 q = sl.v.get(s);
                        xv = (t == 0) ? q.x : (t == 1) ? q.y : (t == 2) ? q.z : q.w;
                    } else {
                        const float4 q = sl.v.get(t);
                        xv = (s == 0) ? q.x : (s == 1) ? q.y : (s == 2) ? q.z : q.w;
                    }
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[t][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(xv, kr[nb][s], acc[t][nb], 0, 0, 0);
                }
            }
        } else {
            bf16x8 krb[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const float4 p0 = sl.pan[nb][0];
                const float4 p1 = sl.pan[nb][PV - 1];
                const float ww = sl.w[nb];
                uint4 u;
                u.x = pack_bf16(p0.x * ww, p0.y * ww);
                u.y = pack_bf16(p0.z * ww, p0.w * ww);
                u.z = pack_bf16(p1.x * ww, p1.y * ww);
                u.w = pack_bf16(p1.z * ww, p1.w * ww);
                krb[nb] = __builtin_bit_cast(bf16x8, u);
            }
            if constexpr (CA) {
                uint4 v[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if constexpr (ST == 1) v[q] = sl.v.get(q);
                    else v[q] = fp8x8_to_bf16x8(sl.v.get(q));
                }
                // in-register 8x8 transpose of 16-bit elements: xa[t] = (v[0][t], ..., v[7][t])
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    uint4 o;
                    o.x = half_of(v[0], t) | (half_of(v[1], t) << 16);
                    o.y = half_of(v[2], t) | (half_of(v[3], t) << 16);
                    o.z = half_of(v[4], t) | (half_of(v[5], t) << 16);
                    o.w = half_of(v[6], t) | (half_of(v[7], t) << 16);
                    const bf16x8 av = __builtin_bit_cast(bf16x8, o);
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[t][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, krb[nb], acc[t][nb], 0, 0, 0);
                }
            } else {
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    uint4 xa;
                    if constexpr (ST == 1) xa = sl.v.get(t);
                    else xa = fp8x8_to_bf16x8(sl.v.get(t));
                    const bf16x8 av = __builtin_bit_cast(bf16x8, xa);
#pragma unroll
                    for (int nb = 0; nb < NB; ++nb)
                        acc[t][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, krb[nb], acc[t][nb], 0, 0, 0);
                }
            }
        }
    };
    // Every issue is unconditional: a load issued on only some paths makes the compiler's
    // wait counting assume the shortest queue and drain all loads in flight.
    Slot ring[PD];
    if (nsteps > 0) {
#pragma unroll
        for (int p = 0; p < PD - 1; ++p) issue(ring[p]);
        for (long base = 0; base < nsteps; base += PD) {
#pragma unroll
            for (int p = 0; p < PD; ++p) {
                issue(ring[(p + PD - 1) % PD]);
                if (base + p < nsteps) compute(ring[p]);
            }
        }
    }

    __syncthreads();  // all waves are done with the weight table that shares the buffer
    // ---- fixed-order reduction of the 4 waves in LDS ----
    // acc[t][nb][v] is MFMA output (row mr = 4*(lane/16)+v, col 16nb + lane%16)
    for (int wv = 0; wv < 4; ++wv) {
        if (wave == wv) {
#pragma unroll
            for (int t = 0; t < T; ++t)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int mr = 4 * lg + v;
                        int lr;
                        if constexpr (CA) lr = (ST == 0) ? (4 * mr + t) : (8 * mr + t);
                        else lr = 16 * t + mr;
                        const int idx = lr * Rp + 16 * nb + lm;
                        red[idx] = (wv == 0 ? 0.0f : red[idx]) + acc[t][nb][v];
                    }
        }
        __syncthreads();
    }
    // zero any padded columns beyond 16*NB (Rp == 16*NB always, so none), write partial
    float *dst = a.partial + (split * a.rows_pad + row0) * Rp;
    for (int e = threadIdx.x; e < RW * Rp; e += 256) {
        float v = red[e];
        if constexpr (CA && ST == 2) {
            const long r = row0 + e / Rp;
            v *= (r < a.In) ? a.row_scale[r] : 0.0f;
        }
        dst[e] = v;
    }
}

// Generic fallback for shapes the register-tile kernel does not cover (odd leading
// extent, rank > 64): one workgroup per (output row, split), one lane per column.
template <int ST>
__global__ __launch_bounds__(64) void k_mttkrp_generic(KArgs a, long kelems) {
    if (a.skip && *a.skip) return;
    const long row = blockIdx.x % a.In;
    const long split = blockIdx.x / a.In;
    const long k0 = split * a.kper, k1 = min(k0 + a.kper, kelems);
    const int Rp = a.Rp;
    for (int c0 = 0; c0 < Rp; c0 += 64) {
        const int r = c0 + threadIdx.x;
        double accd = 0.0;
        for (long e = k0; e < k1; ++e) {
            long rem = e, off = row * a.stride[a.mode];
            float kr = 1.0f;
            for (int m = 0; m < a.order; ++m) {
                if (m == a.mode) continue;
                const long idx = rem % a.dims[m];
                rem /= a.dims[m];
                off += idx * a.stride[m];
                if (r < Rp) kr *= a.panel[m][((idx >> 3) * Rp + r) * 8 + (idx & 7)];
            }
            float x;
            if constexpr (ST == 0) x = reinterpret_cast<const float *>(a.X)[off];
            else if constexpr (ST == 1) x = __uint_as_float((uint32_t)reinterpret_cast<const uint16_t *>(a.X)[off] << 16);
            else {
                const uint8_t q = reinterpret_cast<const uint8_t *>(a.X)[off];
                f32x2 d = __builtin_amdgcn_cvt_pk_f32_fp8((int)q, false);
                x = d.x;
            }
            accd += (double)x * (double)kr;
        }
        if (r < Rp) {
            float v = (float)accd;
            if (ST == 2 && a.mode == 0) v *= a.row_scale[row];
            a.partial[(split * a.rows_pad + row) * Rp + r] = v;
        }
    }
}

__global__ void k_pack_panel(const float *__restrict__ U, long I, long R, int Rp, const float *__restrict__ scale,
                             float *__restrict__ panel, long Ipad) {
    const long n = Ipad * Rp;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long i8 = e & 7;
        const long r = (e >> 3) % Rp;
        const long ib = (e >> 3) / Rp;
        const long i = ib * 8 + i8;
        float v = 0.0f;
        if (i < I && r < R) {
            v = U[i + r * I];
            if (scale) v *= scale[i];
        }
        panel[e] = v;
    }
}

__global__ void k_mttkrp_reduce(const float *__restrict__ partial, float *__restrict__ out, long In, long R, int Rp,
                                long rows_pad, int nsplit) {
    const long n = In * R;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long i = e % In, r = e / In;
        double s = 0.0;
        for (int sp = 0; sp < nsplit; ++sp) s += (double)partial[(sp * rows_pad + i) * Rp + r];
        out[e] = (float)s;
    }
}

// ---------------------------------------------------------------------------
// TDX_MTTKRP_LDS=0 selects the register-tile kernel for every mode (A/B comparisons)
static bool lds_enabled() {
    static int on = -1;
    if (on < 0) {
        const char *e = getenv("TDX_MTTKRP_LDS");
        on = (e && e[0] == '0') ? 0 : 1;
    }
    return on == 1;
}

static int storage_type(int dtype) {
    return dtype == TDX_F32 ? 0 : dtype == TDX_BF16 ? 1 : dtype == TDX_FP8_E4M3 ? 2 : -1;
}

int mttkrp_plan(MttkrpPlan &p, int order, const long *dims, int mode, int dtype, long rank,
                const void *X, int target_wgs, bool split) {
    if (order < 2 || order > 4 || mode < 0 || mode >= order || rank <= 0) return record(TDX_ERR_ARG);
    const int st = storage_type(dtype);
    if (st < 0) return record(TDX_ERR_UNSUPPORTED);
    p = MttkrpPlan();
    p.order = order;
    for (int m = 0; m < order; ++m) {
        if (dims[m] <= 0) return record(TDX_ERR_ARG);
        p.dims[m] = dims[m];
    }
    p.mode = mode;
    p.dtype = dtype;
    p.rank = rank;
    p.Rp = (int)(ceil_div(rank, 16) * 16);
    const long In = dims[mode];
    const int fmode = (mode == 0) ? 1 : 0;
    const int vec = (st == 0) ? 4 : 8;
    const size_t align = (st == 2) ? 8 : 16;
    p.fast = rank <= 64 && (dims[0] % vec == 0) && (((uintptr_t)X) % align == 0);
    if (target_wgs <= 0) target_wgs = 512;
    // the LDS-staged kernel: for output modes > 0 when mode 0 holds whole 16-byte vectors,
    // for mode 0 when it holds whole row groups
    const long lvec = (st == 0) ? 4 : (st == 1) ? 8 : 16;
    if (lds_enabled() && p.fast && mttkrp_lds_covers(order, mode, rank, In, dtype) &&
        (mode == 0 || dims[0] % lvec == 0) && ((uintptr_t)X) % 16 == 0) {
        p.lds = true;
        p.rows_per_wg = 128;
        p.rowtiles = (int)ceil_div(In, 128);
        p.rows_pad = (long)p.rowtiles * 128;
        long S = 1;
        for (int m = 1; m < order; ++m)
            if (m != mode && m != fmode) S *= dims[m];
        p.ksteps = S;
        // grid (measured on the 1024^3 rank-64 tensor, 256-4096 workgroups): one workgroup
        // per CU for BF16, two for FP32, four for FP8 (its tiles carry half the bytes per
        // step).  With the split Khatri-Rao block a step does twice the MFMAs and block work,
        // and one workgroup per CU (one wave per SIMD) left the phases of a step unoverlapped:
        // two per CU ran modes 1 and 2 in 389 us instead of 488 and mode 0 in 438 instead of
        // 485, where the plain BF16 block went from 339 to 360.  A split is one fast-index
        // chunk (or a group of chunks when there are more chunks than splits) times a range of
        // slow indices, so a workgroup loads its factor-panel chunk once and runs one
        // pipelined loop over its slow range.
        const long lds_wgs = (long)device_cus() * ((st == 1) ? (split ? 2 : 1) : (st == 0) ? 2 : 4);
        const long kb = (st == 0) ? 64 : 128;
        const long nfc = ceil_div(dims[fmode], kb);
        long ns = ceil_div(lds_wgs, (long)p.rowtiles);
        if (ns < 1) ns = 1;
        long cps = 1, nss = 1;
        if (ns >= nfc) {
            // at least 8 slow indices per split: small tensors otherwise produce many tiny
            // splits whose partials cost more to sum than they save
            nss = ns / nfc;
            const long cap = ceil_div(S, 8L);
            if (nss > cap) nss = cap;
            if (nss < 1) nss = 1;
        } else {
            cps = ceil_div(nfc, ns);
        }
        p.ksteps_per_split = ceil_div(S, nss);
        nss = ceil_div(S, p.ksteps_per_split);
        p.lds_cps = (int)cps;
        p.lds_nss = (int)nss;
        p.nsplit = (int)(ceil_div(nfc, cps) * nss);
        return TDX_OK;
    }
    if (p.fast) {
        const int KS = (st == 0) ? 16 : 32;
        p.rows_per_wg = (st == 0) ? 64 : 128;
        p.rowtiles = (int)ceil_div(In, p.rows_per_wg);
        p.rows_pad = (long)p.rowtiles * p.rows_per_wg;
        long S = 1;
        for (int m = 0; m < order; ++m)
            if (m != mode && m != fmode) S *= dims[m];
        p.ksteps = ceil_div(dims[fmode], KS) * S;
        long want = ceil_div(target_wgs, p.rowtiles);
        long maxs = ceil_div(p.ksteps, 16);  // at least 4 K steps per wave
        long ns = want < maxs ? want : maxs;
        if (ns < 1) ns = 1;
        p.ksteps_per_split = ceil_div(p.ksteps, ns);
        // the slow-mode weight table of a split (<= kper/FB + 2 slow indices) lives in the
        // kernel's LDS reduction buffer of rows_per_wg * 64 floats
        const long FB = ceil_div(dims[fmode], KS);
        const long cap_sl = (long)p.rows_per_wg * 64 / p.Rp;
        const long kmax = (cap_sl - 2) * FB;
        if (p.ksteps_per_split > kmax) p.ksteps_per_split = kmax;
        p.nsplit = (int)ceil_div(p.ksteps, p.ksteps_per_split);
    } else {
        long kelems = 1;
        for (int m = 0; m < order; ++m)
            if (m != mode) kelems *= dims[m];
        p.rows_per_wg = 1;
        p.rowtiles = (int)In;
        p.rows_pad = In;
        p.ksteps = kelems;
        long want = ceil_div(4096, In);
        long maxs = ceil_div(kelems, 64);
        long ns = want < maxs ? want : maxs;
        if (ns < 1) ns = 1;
        p.ksteps_per_split = ceil_div(kelems, ns);
        p.nsplit = (int)ceil_div(kelems, p.ksteps_per_split);
    }
    return TDX_OK;
}

int pack_panel(const float *U, long I, long R, int Rp, const float *scale, float *panel, hipStream_t s) {
    const long Ipad = panel_rows(I);
    const long n = Ipad * Rp;
    long g = ceil_div(n, 256);
    if (g > 4096) g = 4096;
    k_pack_panel<<<(int)g, 256, 0, s>>>(U, I, R, Rp, scale, panel, Ipad);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

template <int ST, bool CA>
static void launch_fast(const MttkrpPlan &p, const KArgs &a, hipStream_t s) {
    const int grid = p.rowtiles * p.nsplit;
    switch (p.Rp / 16) {
        case 1: k_mttkrp<ST, CA, 1><<<grid, 256, 0, s>>>(a); break;
        case 2: k_mttkrp<ST, CA, 2><<<grid, 256, 0, s>>>(a); break;
        case 3: k_mttkrp<ST, CA, 3><<<grid, 256, 0, s>>>(a); break;
        default: k_mttkrp<ST, CA, 4><<<grid, 256, 0, s>>>(a); break;
    }
}

int mttkrp_launch(const MttkrpPlan &p, const void *X, const float *const *panels, const float *row_scale,
                  float *partial, hipStream_t s, const int *skip) {
    KArgs a{};
    a.skip = skip;
    a.X = X;
    for (int m = 0; m < 4; ++m) a.panel[m] = (m < p.order) ? panels[m] : nullptr;
    a.row_scale = row_scale;
    a.partial = partial;
    long st = 1;
    for (int m = 0; m < 4; ++m) {
        a.dims[m] = (m < p.order) ? p.dims[m] : 1;
        a.stride[m] = st;
        st *= a.dims[m];
    }
    a.order = p.order;
    a.mode = p.mode;
    a.fmode = (p.mode == 0) ? 1 : 0;
    a.In = p.dims[p.mode];
    a.If = p.dims[a.fmode];
    a.Rp = p.Rp;
    a.rowtiles = p.rowtiles;
    a.rows_pad = p.rows_pad;
    a.kper = p.ksteps_per_split;
    a.ksteps = p.ksteps;
    a.nslow = 0;
    for (int m = 0; m < p.order; ++m)
        if (m != p.mode && m != a.fmode) a.slow[a.nslow++] = m;
    const int stt = storage_type(p.dtype);
    if (p.dtype == TDX_FP8_E4M3 && p.mode == 0 && !row_scale) return record(TDX_ERR_ARG);
    if (p.lds) return mttkrp_lds_launch(p, X, panels, row_scale, partial, s, skip);
    if (p.split) return record(TDX_ERR_UNSUPPORTED);  // the split block exists in the LDS kernel only
    if (p.fast) {
        const int KS = (stt == 0) ? 16 : 32;
        a.FB = ceil_div(a.If, KS);
        const bool ca = (p.mode == 0);
        if (stt == 0) { if (ca) launch_fast<0, true>(p, a, s); else launch_fast<0, false>(p, a, s); }
        else if (stt == 1) { if (ca) launch_fast<1, true>(p, a, s); else launch_fast<1, false>(p, a, s); }
        else { if (ca) launch_fast<2, true>(p, a, s); else launch_fast<2, false>(p, a, s); }
    } else {
        a.FB = 1;
        const int grid = (int)(a.In * p.nsplit);
        if (stt == 0) k_mttkrp_generic<0><<<grid, 64, 0, s>>>(a, p.ksteps);
        else if (stt == 1) k_mttkrp_generic<1><<<grid, 64, 0, s>>>(a, p.ksteps);
        else k_mttkrp_generic<2><<<grid, 64, 0, s>>>(a, p.ksteps);
    }
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

int mttkrp_reduce(const MttkrpPlan &p, const float *partial, float *out, hipStream_t s) {
    const long n = p.dims[p.mode] * p.rank;
    long g = ceil_div(n, 256);
    if (g > 4096) g = 4096;
    k_mttkrp_reduce<<<(int)g, 256, 0, s>>>(partial, out, p.dims[p.mode], p.rank, p.Rp, p.rows_pad, p.nsplit);
    TDX_LAUNCH_CHECK();
    return TDX_OK;
}

// ---- scratch of the no-allocation MTTKRP (tdx_mttkrp_ex_noalloc) ----
// One statically allocated device array per GPU, held until the process ends, serves the factor
// panels and the split-K partials.  Host code never passes the array's name as an address (in
// host code it is a stand-in, not the device address): the address is taken once per device,
// with that device current, by hipGetSymbolAddress, and kept per device.  Every launch writes
// all of the scratch it later reads (panels including their zero padding, every partial row of
// every split), so nothing needs clearing between calls.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static constexpr int kScratchDevices = 64;
static void *g_scratch_addr[kScratchDevices];
static std::mutex g_scratch_mutex;

int noalloc_scratch(void **ptr) {
    int dev = 0;
    TDX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) return record(TDX_ERR_UNSUPPORTED);
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return record_hip(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        if (!p) return record(TDX_ERR_HIP);
        g_scratch_addr[dev] = p;
    }
    if (ptr) *ptr = g_scratch_addr[dev];
    return TDX_OK;
}

}  // namespace tdx

using namespace tdx;

extern "C" {

int tdx_mttkrp_ex_noalloc(const void *X, int dtype, const float *scales, int order, const long *dims,
                          const float *const *factors, long rank, int mode, float *out) {
    if (!X || !dims || !factors || !out) return record(TDX_ERR_ARG);
    MttkrpPlan p;
    int st = mttkrp_plan(p, order, dims, mode, dtype, rank, X);
    if (st) return st;
    for (int m = 0; m < order; ++m)
        if (m != mode && !factors[m]) return record(TDX_ERR_ARG);
    if (dtype == TDX_FP8_E4M3 && !scales) return record(TDX_ERR_ARG);
    // scratch layout: the panels of the non-output modes, then the partials (256-byte aligned)
    size_t off[4] = {0, 0, 0, 0}, need = 0;
    for (int m = 0; m < order; ++m) {
        if (m == mode) continue;
        off[m] = need;
        need += (sizeof(float) * (size_t)panel_floats(dims[m], p.Rp) + 255) & ~size_t(255);
    }
    const size_t poff = need;
    need += sizeof(float) * p.partial_floats();
    if (need > kNoAllocScratchBytes) return record(TDX_ERR_TOO_LARGE);
    void *base = nullptr;
    st = noalloc_scratch(&base);
    if (st) return st;
    unsigned char *b = (unsigned char *)base;
    hipStream_t s = stream();
    const float *panels[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int m = 0; m < order; ++m) {
        if (m == mode) continue;
        const float *sc = (dtype == TDX_FP8_E4M3 && m == 0) ? scales : nullptr;
        st = pack_panel(factors[m], dims[m], rank, p.Rp, sc, (float *)(b + off[m]), s);
        if (st) return record(st);
        panels[m] = (const float *)(b + off[m]);
    }
    float *part = (float *)(b + poff);
    st = mttkrp_launch(p, X, panels, (dtype == TDX_FP8_E4M3 && mode == 0) ? scales : nullptr, part, s);
    if (st) return record(st);
    return record(mttkrp_reduce(p, part, out, s));
}

int tdx_mttkrp_ex(const void *X, int dtype, const float *scales, int order, const long *dims,
                  const float *const *factors, long rank, int mode, float *out) {
    if (!X || !dims || !factors || !out) return record(TDX_ERR_ARG);
    MttkrpPlan p;
    int st = mttkrp_plan(p, order, dims, mode, dtype, rank, X);
    if (st) return st;
    for (int m = 0; m < order; ++m)
        if (m != mode && !factors[m]) return record(TDX_ERR_ARG);
    if (dtype == TDX_FP8_E4M3 && !scales) return record(TDX_ERR_ARG);
    hipStream_t s = stream();
    const float *panels[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int m = 0; m < order; ++m) {
        if (m == mode) continue;
        void *ws = nullptr;
        st = workspace(WS_PANEL0 + m, sizeof(float) * panel_floats(dims[m], p.Rp), &ws);
        if (st) return st;
        const float *sc = (dtype == TDX_FP8_E4M3 && m == 0) ? scales : nullptr;
        st = pack_panel(factors[m], dims[m], rank, p.Rp, sc, (float *)ws, s);
        if (st) return record(st);
        panels[m] = (const float *)ws;
    }
    void *part = nullptr;
    st = workspace(WS_PARTIAL, sizeof(float) * p.partial_floats(), &part);
    if (st) return st;
    st = mttkrp_launch(p, X, panels, (dtype == TDX_FP8_E4M3 && mode == 0) ? scales : nullptr, (float *)part, s);
    if (st) return record(st);
    st = mttkrp_reduce(p, (const float *)part, out, s);
    return record(st);
}

// Times `iters` launches of the MTTKRP kernel alone (panels and workspace prepared once)
// with hipEvents on the library stream; *avg_ms receives the mean launch duration.
int tdx_bench_mttkrp(const void *X, int dtype, const float *scales, int order, const long *dims,
                     const float *const *factors, long rank, int mode, int iters, float *avg_ms) {
    if (!X || !dims || !factors || !avg_ms || iters <= 0) return record(TDX_ERR_ARG);
    MttkrpPlan p;
    int st = mttkrp_plan(p, order, dims, mode, dtype, rank, X);
    if (st) return st;
    if (dtype == TDX_FP8_E4M3 && !scales) return record(TDX_ERR_ARG);
    hipStream_t s = stream();
    const float *panels[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int m = 0; m < order; ++m) {
        if (m == mode) continue;
        if (!factors[m]) return record(TDX_ERR_ARG);
        void *ws = nullptr;
        st = workspace(WS_PANEL0 + m, sizeof(float) * panel_floats(dims[m], p.Rp), &ws);
        if (st) return st;
        st = pack_panel(factors[m], dims[m], rank, p.Rp, (dtype == TDX_FP8_E4M3 && m == 0) ? scales : nullptr,
                        (float *)ws, s);
        if (st) return record(st);
        panels[m] = (const float *)ws;
    }
    void *part = nullptr;
    st = workspace(WS_PARTIAL, sizeof(float) * p.partial_floats(), &part);
    if (st) return st;
    const float *rs = (dtype == TDX_FP8_E4M3 && mode == 0) ? scales : nullptr;
    for (int w = 0; w < 2; ++w) {
        st = mttkrp_launch(p, X, panels, rs, (float *)part, s);
        if (st) return record(st);
    }
    hipEvent_t e0, e1;
    TDX_HIP(hipEventCreate(&e0));
    TDX_HIP(hipEventCreate(&e1));
    TDX_HIP(hipEventRecord(e0, s));
    for (int it = 0; it < iters && !st; ++it) st = mttkrp_launch(p, X, panels, rs, (float *)part, s);
    TDX_HIP(hipEventRecord(e1, s));
    TDX_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    TDX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return record(st);
    *avg_ms = ms / (float)iters;
    return record(TDX_OK);
}

// reference mttkrp(X, left, right, XU, flag, m, n, k, r): left/right are the slow/fast
// factors of the Khatri-Rao product (src/mttkrp.cu:11-77)
int tdx_mttkrp(const float *X, const float *left, const float *right, float *XU,
               long flag, long m, long n, long k, long r) {
    if (flag < 1 || flag > 3) return record(TDX_ERR_ARG);
    const long dims[3] = {m, n, k};
    const float *f[3] = {nullptr, nullptr, nullptr};
    if (flag == 1) { f[1] = right; f[2] = left; }
    else if (flag == 2) { f[0] = right; f[2] = left; }
    else { f[0] = right; f[1] = left; }
    return tdx_mttkrp_ex(X, TDX_F32, nullptr, 3, dims, f, r, (int)flag - 1, XU);
}

}  // extern "C"
